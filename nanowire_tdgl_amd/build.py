"""Build libtdglhip.so (the HIP kernels + C ABI) in-tree for gfx950.

    python -m nanowire_tdgl_amd.build

hipcc cross-compiles without a GPU.  The shared library is written next to
this file, where :mod:`nanowire_tdgl_amd._native` loads it.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libtdglhip.so"

# FP64 parity sources: no FMA contraction, so products and sums round like numpy.
SOURCES_EXACT = ["tdgl_f64.hip", "tdgl_api.hip", "tdgl_batch.hip", "tdgl_adjoint.hip",
                 "tdgl_noalloc.hip"]
# Mixed-precision sources (FP32 stencil / Krylov): contraction allowed.
SOURCES_FAST: list[str] = ["tdgl_f32.hip"]

COMMON = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # packed-FP32 instructions off (they have given silent wrong results on these boxes)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build libtdglhip.so)")


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, cwd=str(PKG), capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")


def build(verbose: bool = False) -> Path:
    cc = hipcc()
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    objs = []
    for src, extra in [(s, ["-ffp-contract=off"]) for s in SOURCES_EXACT] + \
                      [(s, ["-ffp-contract=fast"]) for s in SOURCES_FAST]:
        obj = objdir / (Path(src).stem + ".o")
        cmd = [cc, *COMMON, *extra, "-c", str(CSRC / src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        objs.append(str(obj))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [cc, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
