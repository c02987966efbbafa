"""Builds librend.so (host C + gfx950 HIP kernels) in-tree.

No cmake/ninja: each translation unit is one compiler call, run in parallel,
then one hipcc link.  Host C is compiled by gcc with -ffp-contract=off so the
BVH builder and bridge arithmetic round exactly like the reference's parity
build; HIP code is compiled for gfx950 only, with -ffp-contract=off (the
traversal spells out every fused operation) and packed-FP32 instructions
disabled.

    python -m nodebleed_amd.build          # build, print the output path
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
OUT = PKG / "librend.so"
OBJDIR = PKG / "_build"

ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
HIPCC = shutil.which("hipcc") or str(ROCM / "bin" / "hipcc")
ARCH = "gfx950"

HOST_SRC = ["host/bvh8.c", "host/rend.c", "host/frame.c", "host/pt.c",
            "host/scenegen.c", "host/guide.c", "host/approx.c", "host/gltf.c"]
HIP_SRC = ["kern/trace.hip", "kern/frame.hip", "kern/pt.hip", "kern/guide.hip", "kern/tlas.hip",
           "kern/blas.hip"]

CFLAGS = ["-std=c11", "-O2", "-fPIC", "-ffp-contract=off", "-fno-strict-aliasing",
          "-mavx2", "-mfma",
          "-Wall", "-Wextra", "-Wno-unused-parameter",
          "-D__HIP_PLATFORM_AMD__=1", f"-I{ROCM / 'include'}", f"-I{INCLUDE}",
          f"-I{CSRC}", f"-I{CSRC / 'host'}"]
HIPFLAGS = [f"--offload-arch={ARCH}", "-O3", "-fPIC", "-ffp-contract=off",
            "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
            "-packed-fp32-ops", f"-I{INCLUDE}", f"-I{CSRC}", f"-I{CSRC / 'kern'}"]


def _run(cmd: list[str]) -> str:
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{p.stdout}\n{p.stderr}")
    return p.stderr


def _compile(src: str) -> Path:
    obj = OBJDIR / (src.replace("/", "_") + ".o")
    path = CSRC / src
    if src.endswith(".hip"):
        _run([HIPCC, *HIPFLAGS, "-c", str(path), "-o", str(obj)])
    else:
        _run(["gcc", *CFLAGS, "-c", str(path), "-o", str(obj)])
    return obj


def build(verbose: bool = False) -> Path:
    """Compiles every source and links nodebleed_amd/librend.so."""
    OBJDIR.mkdir(exist_ok=True)
    srcs = HOST_SRC + HIP_SRC
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = OUT.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *map(str, objs),
          "-o", str(tmp), "-lpthread", "-lmvec", "-lm"])
    os.replace(tmp, OUT)
    if verbose:
        print(OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
