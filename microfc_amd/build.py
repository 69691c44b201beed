"""Builds the HIP library libmfc_rhs.so for gfx950, in-tree.

Each .hip translation unit is compiled by its own hipcc process (the sweep
kernels are split by number of fluids and by form so the variants build in
parallel), then linked into microfc_amd/libmfc_rhs.so.  Flags that matter
for results:
  -ffp-contract=off   every multiply and add rounds separately, as in the
                      reference's statements (FP64 parity is bit-level);
  -ffp-contract=fast  only for the *_f32.hip units (the FP32-bulk and
                      learned sweeps) and the *_f64fast.hip units (the
                      contracted FP64 sweeps), whose error against the plain
                      FP64 form is measured, not bit-level: products feeding a
                      sum become FMAs;
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed FP32 instructions.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")
LIB = os.path.join(HERE, "libmfc_rhs.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MFC_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-I", os.path.join(HERE, "..", "include"),
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _flags(src: str) -> list[str]:
    if src.endswith("_f32.hip") or src.endswith("_f64fast.hip"):
        return ["-ffp-contract=fast" if f == "-ffp-contract=off" else f for f in FLAGS]
    return FLAGS


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.basename(src)[:-4] + ".o")
    cmd = [HIPCC, *_flags(src), "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


def build(verbose: bool = False, jobs: int | None = None) -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    r = subprocess.run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
