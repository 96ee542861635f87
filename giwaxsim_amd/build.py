"""Builds the gfx950 C-ABI library ``giwaxsim_amd/lib/libgiwaxsim_amd.so``.

hipcc cross-compiles for gfx950 without a GPU.  Flags:
  -ffp-contract=off   every FP64 index operation rounds on its own, like
                      numpy's elementwise evaluation (explicit fma() calls are
                      kept where wanted)
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed FP32 instructions in the generated code
"""

from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libgiwaxsim_amd.so")
SOURCES = ["gx_frames.hip", "gx_detector.hip", "gx_adjoint.hip", "gx_compare.hip", "gx_slab_build.hip",
           "gx_workspace.hip"]
ARCH = os.environ.get("GX_OFFLOAD_ARCH", "gfx950")


def hipcc_command(out=LIB, arch=ARCH):
    return [
        "hipcc", f"--offload-arch={arch}", "-O3", "-std=c++17", "-fPIC", "-shared",
        "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
        "-I", os.path.join(HERE, "..", "include"),
        "-o", out,
    ] + [os.path.join(CSRC, s) for s in SOURCES]


def build(verbose=False):
    os.makedirs(LIBDIR, exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = hipcc_command(out=tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libgiwaxsim_amd.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
