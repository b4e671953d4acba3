"""CPU: the FP32 sweeps' branch-light cell staging (tdgl_stage32.hpp) equals
the reference-ordered boundary rules of eff_cell_pair (tdgl_device.hpp, which
the GPU parity tests pin to tdgl3d's _apply_boundary_conditions,
rhs.py:54-139) bit for bit on every cell of 300 random 2-D/3-D grids with
periodic axes, uniform and per-node fields and hole bits.  hipcc compiles the
__host__ __device__ functions into a host program; no GPU is used."""

from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def _hipcc():
    for cand in ("/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    return None


@pytest.mark.skipif(_hipcc() is None, reason="hipcc not installed")
def test_stage32_matches_boundary_rules(tmp_path):
    exe = tmp_path / "stage32_check"
    cmd = [_hipcc(), "-O1", "-std=c++17", f"-I{ROOT / 'include'}",
           f"-I{ROOT / 'nanowire_tdgl_amd' / 'csrc'}", str(ROOT / "tests" / "stage32_check.cpp"),
           "-o", str(exe)]
    build = subprocess.run(cmd, capture_output=True, text=True, cwd=str(tmp_path))
    assert build.returncode == 0, build.stderr[-3000:]
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    print(run.stdout[-2000:])
    assert run.returncode == 0 and run.stdout.strip().startswith("OK"), run.stdout[-2000:]


@pytest.mark.skipif(_hipcc() is None, reason="hipcc not installed")
def test_per_film_sweep_tiles_fit_the_block(tmp_path):
    """sweep_film_tile (tdgl_launch.hpp) on every film from 1 to 400 nodes
    across and along: the tile fits the 256-lane block and its shared memory,
    every lane's LDS cell stays inside the staged cells, and the pick is the
    cost model's argmin."""
    exe = tmp_path / "sweep_tile_check"
    cmd = [_hipcc(), "-O1", "-std=c++17", f"-I{ROOT / 'include'}",
           f"-I{ROOT / 'nanowire_tdgl_amd' / 'csrc'}", str(ROOT / "tests" / "sweep_tile_check.cpp"),
           "-o", str(exe)]
    build = subprocess.run(cmd, capture_output=True, text=True, cwd=str(tmp_path))
    assert build.returncode == 0, build.stderr[-3000:]
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    print(run.stdout[-2000:])
    assert run.returncode == 0 and run.stdout.strip().startswith("OK"), run.stdout[-2000:]
