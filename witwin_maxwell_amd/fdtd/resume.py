"""Detached forward state for resuming a prepared FDTD run.

Mirrors the reference's resume contract (witwin/maxwell/fdtd/resume.py:73-160
``FDTDResumeCheckpoint``, :372-437 ``capture_resume_checkpoint``, :661-700
``preflight_resume_checkpoint`` / ``restore_resume_checkpoint``): a checkpoint
taken after ``step`` steps holds every tensor the remaining steps read -- fields,
psi slabs, pole state, the TFSF auxiliary line -- and every running sum they add
to (observer and full-field DFT accumulators), detached on the host, plus a
signature of the run's layout.  Restoring it into a fresh run of the same plan
and finishing the steps gives the uninterrupted run's results.
"""

from __future__ import annotations

from dataclasses import dataclass
from pathlib import Path
from typing import Any

import torch

from .runtime import DeviceRun

_SCHEMA_VERSION = 1


def _signature(run: DeviceRun) -> dict:
    plan = run.plan
    return {
        "shape": list(plan.shape),
        "dt": float(plan.dt),
        "cpml": bool(plan.uses_cpml),
        "psi": {name: list(t.shape) for name, t in run.psi.items()},
        "poles": {c: len(states) for c, states in run.ade_state.items()},
        "planes": [list(t.shape) for t in run.plane_re],
        "points": None if run.point_re is None else list(run.point_re.shape),
        "dft": {c: list(t.shape) for c, t in run.dft_re.items()},
        "aux": None if run.aux_e is None else int(run.aux_e.numel()),
    }


def _state_tensors(run: DeviceRun) -> dict:
    """name -> live tensor of everything a resumed run needs (state of step
    ``run.step_index``, in the buffers of that step's parity)."""
    out = {f"field/{c}": t for c, t in run.fields.items()}
    for name in run.psi:
        out[f"psi/{name}"] = run.psi_current(name)
    for c, states in run.ade_state.items():
        for i, s in enumerate(states):
            for k, v in s.items():
                if v is not None:
                    out[f"pole/{c}/{i}/{k}"] = v
    for q, (re, im) in enumerate(zip(run.plane_re, run.plane_im)):
        out[f"plane/{q}/re"] = re
        out[f"plane/{q}/im"] = im
    if run.point_re is not None:
        out["point/re"] = run.point_re
        out["point/im"] = run.point_im
    for c in run.dft_re:
        out[f"dft/{c}/re"] = run.dft_re[c]
        out[f"dft/{c}/im"] = run.dft_im[c]
    if run.aux_e is not None:
        out["aux/e"] = run.aux_e
        out["aux/h"] = run.aux_h
    return out


@dataclass(frozen=True)
class FDTDResumeCheckpoint:
    """Detached forward-execution state for resuming a prepared FDTD run."""

    step: int
    total_steps: int
    dt: float
    tensors: dict
    signature: dict

    schema_version = _SCHEMA_VERSION

    def save(self, path: str | Path) -> None:
        out = Path(path)
        out.parent.mkdir(parents=True, exist_ok=True)
        torch.save({"data_type": "FDTDResumeCheckpoint", "schema_version": self.schema_version, "step": self.step,
                    "total_steps": self.total_steps, "dt": self.dt, "tensors": self.tensors,
                    "signature": self.signature}, out)

    @classmethod
    def load(cls, path: str | Path, *, map_location: str | torch.device = "cpu") -> "FDTDResumeCheckpoint":
        payload = torch.load(Path(path), map_location=map_location, weights_only=True)
        if not isinstance(payload, dict) or payload.get("data_type") != "FDTDResumeCheckpoint":
            raise ValueError("FDTD resume checkpoint has an invalid data_type.")
        if payload.get("schema_version") != cls.schema_version:
            raise ValueError(f"Unsupported FDTD resume checkpoint schema_version {payload.get('schema_version')!r}.")
        missing = {"step", "total_steps", "dt", "tensors", "signature"}.difference(payload)
        if missing:
            raise ValueError("FDTD resume checkpoint is missing required keys: " + ", ".join(sorted(missing)))
        return cls(step=int(payload["step"]), total_steps=int(payload["total_steps"]), dt=float(payload["dt"]),
                   tensors=dict(payload["tensors"]), signature=dict(payload["signature"]))


def capture_resume_checkpoint(run: DeviceRun, *, step: int, total_steps: int) -> FDTDResumeCheckpoint:
    if not 0 <= int(step) < int(total_steps):
        raise ValueError("Resume checkpoint step must satisfy 0 <= step < total_steps.")
    if run.step_index != int(step):
        raise RuntimeError(f"the run completed {run.step_index} steps, not {step}")
    torch.cuda.synchronize(run.device)
    tensors = {k: v.detach().to("cpu", copy=True) for k, v in _state_tensors(run).items()}
    return FDTDResumeCheckpoint(step=int(step), total_steps=int(total_steps), dt=float(run.plan.dt),
                                tensors=tensors, signature=_signature(run))


def preflight_resume_checkpoint(run: DeviceRun, checkpoint: Any, *, total_steps: int) -> None:
    if not isinstance(checkpoint, FDTDResumeCheckpoint):
        raise TypeError("resume_from must be an FDTDResumeCheckpoint.")
    if checkpoint.total_steps != int(total_steps):
        raise ValueError(f"Resume checkpoint planned {checkpoint.total_steps} steps, "
                         f"but this simulation plans {total_steps}.")
    if not 0 <= checkpoint.step < checkpoint.total_steps:
        raise ValueError("FDTD resume checkpoint step is outside the planned run.")
    if checkpoint.dt != float(run.plan.dt):
        raise ValueError(f"FDTD resume checkpoint dt={checkpoint.dt!r} does not match the prepared run's "
                         f"dt={float(run.plan.dt)!r}.")
    if checkpoint.signature != _signature(run):
        raise ValueError("FDTD resume checkpoint does not match the prepared run's layout.")


def restore_resume_checkpoint(run: DeviceRun, checkpoint: FDTDResumeCheckpoint, *, total_steps: int) -> int:
    """Load ``checkpoint`` into ``run`` (buffers of the checkpoint step's parity);
    returns the step to continue from."""
    preflight_resume_checkpoint(run, checkpoint, total_steps=total_steps)
    run.reset()
    run.step_index = checkpoint.step
    live = _state_tensors(run)
    if set(live) != set(checkpoint.tensors):
        raise ValueError("FDTD resume checkpoint does not hold the prepared run's state tensors.")
    for k, t in live.items():
        t.copy_(checkpoint.tensors[k])
    return checkpoint.step


__all__ = ["FDTDResumeCheckpoint", "capture_resume_checkpoint", "preflight_resume_checkpoint",
           "restore_resume_checkpoint"]
