"""Builds twixtor_amd/libtwx.so, the HIP backend for gfx950, in-tree.

Each source is compiled to its own object in parallel (no kernel calls cross a translation unit,
so no relocatable device code is needed), then the objects are linked into the shared library.

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
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libtwx.so")
OBJDIR = os.path.join(HERE, "_build")
SOURCES = ["twx_device.hip", "twx_api.hip", "twx_selfcheck.hip", "twx_diff.hip", "twx_forward.hip"]
ARCH = os.environ.get("TWX_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    return subprocess.run(cmd, capture_output=True, text=True)


def build(verbose: bool = False) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    os.makedirs(OBJDIR, exist_ok=True)
    objs = [os.path.join(OBJDIR, os.path.splitext(s)[0] + ".o") for s in SOURCES]
    cmds = [[hipcc, *FLAGS, "-c", "-o", o, os.path.join(CSRC, s)] for s, o in zip(SOURCES, objs)]
    with ThreadPoolExecutor(max_workers=min(len(cmds), os.cpu_count() or 1)) as ex:
        results = list(ex.map(lambda c: _run(c, verbose), cmds))
    failed = [r for r in results if r.returncode != 0]
    for r in results:
        if r.returncode != 0 or (verbose and r.stderr):
            sys.stderr.write(r.stdout + r.stderr)
    if failed:
        raise RuntimeError("hipcc failed compiling libtwx.so's sources")
    tmp = LIB + ".tmp"
    res = _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs], verbose)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed linking libtwx.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
