"""Builds libseqan_amd.so in-tree with hipcc for gfx950.

    python -m seqan_amd.build            # or: python seqan_amd/build.py

The shared library holds the HIP kernels, the host engine behind the C ABI
(include/seqan_amd.h), the scoring presets and the synthetic workload
generator.  It is built next to this file so that it travels with the tree.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libseqan_amd.so")
ARCH = os.environ.get("SQMI_OFFLOAD_ARCH", "gfx950")

SOURCES = ["kernels.hip", "long.hip", "banded.hip", "engine.cpp", "scoring.cpp", "workload.cpp"]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise FileNotFoundError("hipcc not found")


def build(verbose: bool = False) -> str:
    cc = hipcc()
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    common = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result",
              f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}"]
    objs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(objdir, src + ".o")
        cmd = [cc] + common
        if src.endswith(".hip"):
            # Device code: gfx950 only; packed FP32 instructions disabled.
            cmd += [f"--offload-arch={ARCH}", "-x", "hip",
                    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
        else:
            cmd += ["-x", "c++", "-fopenmp", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"]
        cmd += ["-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"compiling {src} failed")
        objs.append(obj)
    tmp = LIB + ".tmp"
    link = [cc, "-shared", "-fPIC", "-fopenmp", f"--offload-arch={ARCH}", "-o", tmp] + objs + [
        "-Wl,-rpath,/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib/llvm/lib", "-Wl,-rpath,/opt/rocm/llvm/lib"]
    if verbose:
        print(" ".join(link), flush=True)
    r = subprocess.run(link, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("linking libseqan_amd.so failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
