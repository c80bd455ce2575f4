"""Builds libpe_amd.so (HIP, gfx950) in-tree.

    python physics-engine_amd/build.py

Flags: no FMA contraction and correctly rounded FP32 divide/sqrt (the kernels restate the
reference's FP32 arithmetic operation by operation), denormals kept, packed-FP32 ops off.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SOURCES = ["csrc/pe_world.hip", "csrc/pe_batch.hip", "csrc/pe_host.hip", "csrc/pe_noalloc.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
         "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-gpu-flush-denormals-to-zero",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-value", "-Wno-unused-result"]


def build(verbose: bool = False) -> str:
    out = os.path.join(HERE, "lib", "libpe_amd.so")
    os.makedirs(os.path.dirname(out), exist_ok=True)
    srcs = [os.path.join(HERE, s) for s in SOURCES if os.path.exists(os.path.join(HERE, s))]
    tmp = out + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-o", tmp] + srcs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed: " + " ".join(cmd))
    if verbose:
        sys.stderr.write(r.stderr)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
