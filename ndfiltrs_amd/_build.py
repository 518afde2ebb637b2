"""Compile the HIP library ``libndfgpu.so`` for gfx950 (in-tree)."""

from __future__ import annotations

import os
import shutil
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBRARY = os.path.join(HERE, "libndfgpu.so")
SOURCES = ("ndf_host.hip", "ndf_generic.hip", "ndf_slide.hip", "ndf_bracket.hip", "ndf_network.hip",
           "ndf_stack.hip", "ndf_extrema.hip", "ndf_mesh.hip", "ndf_transfer.hip")

# Strict IEEE for the f64 midpoint (no contraction, denormals kept), packed
# FP32 instructions disabled, gfx950 only.
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
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


def _report(result: subprocess.CompletedProcess) -> None:
    # The host pass does not know the packed-fp32 feature; that note is noise.
    noise = "is not a recognized feature for this target"
    for line in (result.stdout + result.stderr).splitlines():
        if line.strip() and noise not in line:
            print(line, flush=True)


def build(verbose: bool = False) -> str:
    """Compile every HIP source into ``ndfiltrs_amd/libndfgpu.so``: one hipcc
    per source in parallel, then one link."""
    tmp = LIBRARY + ".tmp"
    with tempfile.TemporaryDirectory(prefix="ndfgpu_build_") as objdir:
        jobs = []
        for source in SOURCES:
            obj = os.path.join(objdir, os.path.splitext(source)[0] + ".o")
            cmd = [hipcc(), *FLAGS, "-c", "-o", obj, os.path.join(CSRC, source)]
            if verbose:
                print(" ".join(cmd), flush=True)
            jobs.append((obj, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
        failed = []
        for obj, proc in jobs:
            out, err = proc.communicate()
            _report(subprocess.CompletedProcess(proc.args, proc.returncode, out, err))
            if proc.returncode != 0:
                failed.append((os.path.basename(obj), proc.returncode))
        if failed:
            raise RuntimeError(f"hipcc failed: {failed}")
        cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *(obj for obj, _ in jobs)]
        if verbose:
            print(" ".join(cmd), flush=True)
        result = subprocess.run(cmd, capture_output=True, text=True)
        _report(result)
        if result.returncode != 0:
            raise RuntimeError(f"hipcc link failed with exit code {result.returncode}")
    os.replace(tmp, LIBRARY)
    return LIBRARY


if __name__ == "__main__":
    print(build(verbose=True))
