"""Compile the HIP library ``libndfgpu.so`` for gfx950 (in-tree)."""

from __future__ import annotations

import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBRARY = os.path.join(HERE, "libndfgpu.so")
SOURCES = ("ndf_host.hip", "ndf_generic.hip", "ndf_slide.hip", "ndf_bracket.hip", "ndf_network.hip")

# Strict IEEE for the f64 midpoint (no contraction, denormals kept), packed
# FP32 instructions disabled, gfx950 only.
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wall",
]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found; libndfgpu.so cannot be built")
    return path


def build(verbose: bool = False) -> str:
    """Compile every HIP source into ``ndfiltrs_amd/libndfgpu.so``."""
    tmp = LIBRARY + ".tmp"
    cmd = [hipcc(), *FLAGS, "-o", tmp, *(os.path.join(CSRC, s) for s in SOURCES)]
    if verbose:
        print(" ".join(cmd), flush=True)
    result = subprocess.run(cmd, capture_output=True, text=True)
    # The host pass does not know the packed-fp32 feature; that note is noise.
    noise = "is not a recognized feature for this target"
    for line in (result.stdout + result.stderr).splitlines():
        if line.strip() and noise not in line:
            print(line, flush=True)
    if result.returncode != 0:
        raise RuntimeError(f"hipcc failed with exit code {result.returncode}")
    os.replace(tmp, LIBRARY)
    return LIBRARY


if __name__ == "__main__":
    print(build(verbose=True))
