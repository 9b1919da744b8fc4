"""Builds twixtor_amd/libtwx.so, the HIP backend for gfx950, in-tree.

Flags that the bit-exact contract depends on:
  -ffp-contract=off                       no fused multiply-add: every product and sum rounds on its own
  -fhip-fp32-correctly-rounded-divide-sqrt   IEEE division (the CPU's divss)
  -fno-gpu-flush-denormals-to-zero        f32 denormals preserved
  -target-feature -packed-fp32-ops        no packed FP32 instructions
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libtwx.so")
SOURCES = ["twx_device.hip", "twx_api.hip", "twx_selfcheck.hip", "twx_diff.hip", "twx_forward.hip"]
ARCH = os.environ.get("TWX_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def build(verbose: bool = False) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = LIB + ".tmp"
    cmd = [hipcc, *FLAGS, "-o", tmp] + [os.path.join(CSRC, s) for s in SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libtwx.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
