"""The C-ABI library loads on the CPU host and exports every entry point that
include/tdgl_hip.h declares (no compute calls: there is no GPU here)."""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import pytest

from nanowire_tdgl_amd import _native

ROOT = Path(__file__).resolve().parent.parent


def _declared() -> list[str]:
    names = []
    for h in sorted((ROOT / "include").glob("*.h")):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*(?:int|int64_t|const char\s*\*|void)\s+(tdgl_\w+)\s*\(", text, re.M)
    return names


def test_header_declares_entry_points():
    assert len(_declared()) >= 15


@pytest.mark.skipif(not _native.LIB_PATH.exists(), reason="libtdglhip.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_native.LIB_PATH))
    missing = [s for s in _declared() if not hasattr(lib, s)]
    assert not missing, missing
    assert set(_native.EXPORTS) == set(_declared())


@pytest.mark.skipif(not _native.LIB_PATH.exists(), reason="libtdglhip.so not built")
def test_loader_sets_signatures():
    lib = _native.load()
    assert lib.tdgl_ctx_create.restype is ctypes.c_int
    assert ctypes.sizeof(_native.StepStats) == 8 * 4 + 5 * 8
    text = (ROOT / "include" / "tdgl_hip.h").read_text()
    assert len(_native.KERNEL_CLASSES) == int(re.search(r"#define TDGL_K_COUNT (\d+)", text).group(1))
    assert ctypes.sizeof(_native.SolverOpts) == 8 + 8 + 4 + 4 + 8 + 8 + 8 + 8 + 4 + 4 + 8 + 8


def test_precond_option_codes():
    assert _native.precond_code("chebyshev") == -1
    assert _native.precond_code("none") == 0 and _native.precond_code(None) == 0
    assert _native.precond_code(7) == 7
    for bad in (0, -2, True, "cheb", 2.5):
        with pytest.raises(ValueError):
            _native.precond_code(bad)
    assert _native.solver_opts(precision="mixed").mixed_precond == -1


def test_struct_layout_matches_header():
    text = (ROOT / "include" / "tdgl_hip.h").read_text()
    for st, fields in (("tdgl_grid", _native.Grid._fields_),
                       ("tdgl_solver_opts", _native.SolverOpts._fields_),
                       ("tdgl_step_stats", _native.StepStats._fields_)):
        body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (st, st), text, re.S).group(1)
        body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
        declared = re.findall(r"(\w+)\s*(?:,|;)", body)
        assert [f for f, _ in fields] == declared, st


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_native, "LIB_PATH", tmp_path / "libtdglhip.so")
    monkeypatch.setattr(_native, "_lib", None)
    with pytest.raises(ImportError, match="build it"):
        _native.load()


def test_device_vector_check_rejects_host_and_strided_tensors():
    """The bindings hand kernels raw pointers; _is_vec admits only contiguous
    device tensors of the right dtype and length (checked here without a GPU:
    every host tensor fails it)."""
    import torch
    from nanowire_tdgl_amd._native import _is_vec
    x = torch.zeros(10, dtype=torch.float64)
    assert not _is_vec(x, torch.float64, 10)          # host memory
    assert not _is_vec(x[::2], torch.float64, 5)      # strided
    assert not _is_vec(x.float(), torch.float64, 10)  # dtype
    assert not _is_vec(x, torch.float64, 9)           # length
    assert not _is_vec(None, torch.float64, 10)


def test_solver_opts_preconditioner_codes():
    """solver_opts: the mixed method's preconditioner defaults to the automatic
    Chebyshev degree, the FP64 method's to none (the reference's TGCR)."""
    from nanowire_tdgl_amd import _native
    o = _native.solver_opts()
    assert (o.mixed_precond, o.f64_precond) == (-1, 0)
    o = _native.solver_opts(precision="f64", f64_precond="chebyshev", precond="none")
    assert (o.mixed_precond, o.f64_precond) == (0, -1)
    assert _native.solver_opts(f64_precond=5).f64_precond == 5
