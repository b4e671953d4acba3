"""Run metadata written by solve() (mirror of tdgl3d's io/logging.py:21-314:
same record fields, JSON file per run)."""

from __future__ import annotations

import json
import platform
import subprocess
import time
from dataclasses import asdict, dataclass, field
from datetime import datetime
from pathlib import Path
from typing import Any, Optional

import numpy as np


@dataclass
class RunMetadata:
    timestamp: str
    wall_time_seconds: float
    hostname: str
    platform_info: dict
    git_commit: Optional[str]
    parameters: dict
    device_config: dict
    solver_config: dict
    convergence_info: Optional[dict] = None
    vortex_count_final: Optional[int] = None
    performance_metrics: dict = field(default_factory=dict)

    def to_dict(self) -> dict:
        return asdict(self)

    @classmethod
    def from_dict(cls, data: dict) -> "RunMetadata":
        return cls(**data)

    def save_json(self, filepath) -> None:
        p = Path(filepath)
        p.parent.mkdir(parents=True, exist_ok=True)
        p.write_text(json.dumps(self.to_dict(), indent=2, default=_json_default))

    @classmethod
    def load_json(cls, filepath) -> "RunMetadata":
        return cls.from_dict(json.loads(Path(filepath).read_text()))


def _json_default(obj: Any):
    if isinstance(obj, (np.integer, np.floating)):
        return obj.item()
    if isinstance(obj, np.ndarray):
        return obj.tolist()
    raise TypeError(f"Object of type {type(obj)} is not JSON serializable")


def _git_commit() -> Optional[str]:
    try:
        r = subprocess.run(["git", "rev-parse", "HEAD"], capture_output=True, text=True,
                           timeout=1.0, check=False)
        return r.stdout.strip() if r.returncode == 0 else None
    except (FileNotFoundError, subprocess.TimeoutExpired):
        return None


def create_run_metadata(params, device, method: str, dt: float, t_final: float, wall_time: float,
                        atol: Optional[float] = None, rtol: Optional[float] = None,
                        **extra) -> RunMetadata:
    """Record of a run (logging.py:117-235)."""
    plat = {"os": platform.system(), "os_version": platform.version(),
            "python_version": platform.python_version(), "machine": platform.machine(),
            "processor": platform.processor()}
    pdict = {k: getattr(params, k) for k in ("Nx", "Ny", "Nz", "hx", "hy", "hz", "kappa",
                                              "periodic_x", "periodic_y", "periodic_z")}
    dev = {"has_trilayer": device.trilayer is not None}
    if device.trilayer is not None:
        tr = device.trilayer
        dev["trilayer"] = {"top_thickness": tr.top.thickness_z, "top_kappa": tr.top.kappa,
                           "insulator_thickness": tr.insulator.thickness_z,
                           "bottom_thickness": tr.bottom.thickness_z,
                           "bottom_kappa": tr.bottom.kappa}
    f = device.applied_field
    if f is not None:
        if callable(f):
            dev["applied_field"] = {"type": "callable", "description": str(f)}
        else:
            dev["applied_field"] = {"type": type(f).__name__, "Bx": f.Bx, "By": f.By, "Bz": f.Bz,
                                    "B_initial": f.evaluate(0.0, 1.0)}
    solver = {"method": method, "dt_initial": dt, "t_final": t_final}
    if atol is not None:
        solver["atol"] = atol
    if rtol is not None:
        solver["rtol"] = rtol
    perf = {"wall_time_per_timestep_avg": extra.get("wall_time_per_step"),
            "total_timesteps": extra.get("total_steps"),
            "newton_iterations_total": extra.get("newton_iterations"),
            "dt_reductions": extra.get("dt_reductions")}
    perf = {k: v for k, v in perf.items() if v is not None}
    perf.update({k: v for k, v in extra.items() if k not in perf})
    return RunMetadata(timestamp=datetime.now().isoformat(), wall_time_seconds=wall_time,
                       hostname=platform.node(), platform_info=plat, git_commit=_git_commit(),
                       parameters=pdict, device_config=dev, solver_config=solver,
                       performance_metrics=perf)


class TimingContext:
    def __init__(self):
        self.start_time: Optional[float] = None
        self.end_time: Optional[float] = None

    def __enter__(self) -> "TimingContext":
        self.start_time = time.perf_counter()
        return self

    def __exit__(self, *exc):
        self.end_time = time.perf_counter()
        return False

    @property
    def elapsed(self) -> float:
        if self.start_time is None:
            raise RuntimeError("Timer not started")
        end = self.end_time if self.end_time is not None else time.perf_counter()
        return end - self.start_time
