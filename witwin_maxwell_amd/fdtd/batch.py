"""Many independent scenes advancing in one launch per step.

``BatchRun`` wraps ``mw_batch_*``: every scene is a single-pass ``DeviceRun`` (its
buffers stay in the run), and one launch per kernel of the step covers the tiles
of all scenes, with the step index in a device counter so a run of steps is
captured once as a hipGraph and replayed.  ``autotune`` picks the x-planes per
tile and the number of steps per graph by timing candidates on the batch itself
(then resetting it), and caches the choice per scene-set signature.

The reference has no batched single-GPU path: it runs one solver per scene and
spreads whole solvers over GPUs (execution/ensemble.py:166-180).
"""

from __future__ import annotations

import ctypes
import json
import os

import torch

from . import native as nat
from .runtime import DeviceRun

_CACHE_ENV = "WITWIN_MAXWELL_AMD_TUNE_CACHE"


def _signature(runs):
    parts = []
    for r in runs:
        p = r.plan
        parts.append(f"{p.shape}|{int(p.uses_cpml)}|{int(p.dispersive)}|{len(p.dft_frequencies)}")
    return ";".join(parts)


class BatchRun:
    def __init__(self, runs, xchunk: int = 0):
        if not runs:
            raise ValueError("a batch needs at least one scene")
        self.runs = list(runs)
        for r in self.runs:
            if not r.fused:
                raise ValueError("batched scenes must use the single-pass step (no periodic faces)")
        self.lib = self.runs[0].lib
        self.xchunk = int(xchunk)
        self._handle = None
        self._create()
        self.step_index = 0
        self.graph_steps = 1

    def _create(self):
        progs = (ctypes.POINTER(nat.MwProgram) * len(self.runs))(
            *[ctypes.pointer(r.program) for r in self.runs])
        handle = ctypes.c_void_p()
        with torch.cuda.device(self.runs[0].device):
            rc = self.lib.mw_batch_create(progs, len(self.runs), self.xchunk, ctypes.byref(handle))
        nat.check(rc, self.lib)
        self._progs = progs
        self._handle = handle

    def close(self):
        if self._handle is not None and self._handle.value:
            self.lib.mw_batch_destroy(self._handle)
        self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def total_steps(self):
        return min(r.plan.time_steps for r in self.runs)

    def run(self, count=None, graph_steps=None):
        count = self.total_steps - self.step_index if count is None else int(count)
        if count <= 0:
            return self
        g = self.graph_steps if graph_steps is None else int(graph_steps)
        with torch.cuda.device(self.runs[0].device):
            rc = self.lib.mw_batch_run(self._handle, self.step_index, count, g, nat.current_stream())
        nat.check(rc, self.lib)
        self.step_index += count
        for r in self.runs:
            r.step_index = self.step_index
        return self

    def reset(self):
        for r in self.runs:
            r.reset()
        self.step_index = 0

    # ------------------------------------------------------------ tuning
    def autotune(self, trial_steps: int = 16, chunks=(0, 8, 16, 32), graphs=(1, 8, 32), cache=True):
        """Pick (xchunk, graph_steps) by timing ``trial_steps`` steps of each
        candidate on this batch, then reset every scene."""
        sig = _signature(self.runs)
        path = os.environ.get(_CACHE_ENV)
        table = {}
        if cache and path and os.path.exists(path):
            with open(path) as f:
                table = json.load(f)
        if sig in table:
            self._apply(*table[sig])
            return tuple(table[sig])
        best, best_t = None, float("inf")
        for chunk in chunks:
            self._apply(chunk, 1)
            for g in graphs:
                steps = max(trial_steps, 2 * g)
                if steps > self.total_steps:
                    continue
                self.reset()
                self.run(g if g > 1 else 1, graph_steps=g)  # warm-up incl. graph capture
                torch.cuda.synchronize()
                t0 = torch.cuda.Event(enable_timing=True)
                t1 = torch.cuda.Event(enable_timing=True)
                t0.record()
                self.run(steps, graph_steps=g)
                t1.record()
                torch.cuda.synchronize()
                per_step = t0.elapsed_time(t1) / steps
                if per_step < best_t:
                    best, best_t = (chunk, g), per_step
        self.reset()
        self._apply(*best)
        if cache and path:
            table[sig] = list(best)
            with open(path, "w") as f:
                json.dump(table, f)
        return best

    def _apply(self, chunk, graph_steps):
        if int(chunk) != self.xchunk:
            self.close()
            self.xchunk = int(chunk)
            self._create()
        self.graph_steps = int(graph_steps)
