"""Builds libwavebem_bem.so (HIP, gfx950) in-tree with hipcc.

Usage: python -m wavebem_amd.build   (or wavebem_amd.build.build())
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libwavebem_bem.so")
SOURCES = ["wb_setup.cpp", "wb_host.cpp", "wb_assemble.hip", "wb_adjoint.hip", "wb_project.hip",
           "wb_gemv.hip", "wb_multi.hip", "wb_krylov.hip", "wb_linalg.hip", "wb_coarse.hip",
           "wb_hmat.hip", "wb_noalloc.hip",
           "wb_context.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("WAVEBEM_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # keep the compiler from forming packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # let both 16-double register tiles of the assembly kernel (G and H, 256 bytes) live
    # in VGPRs; the default budget keeps only one and sends the other to scratch
    "-mllvm", "-amdgpu-promote-alloca-to-vector-limit=1024",
    "-Wall",
    "-Wno-unused-result",
    "-Wno-unused-variable",
]


def build(verbose: bool = False, extra=None) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + FLAGS + (extra or []) + ["-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libwavebem_bem.so")
    if verbose and r.stderr.strip():
        sys.stderr.write(r.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
