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
                raise ValueError("batched scenes must use the single-pass step")
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
    def autotune(self, trial_steps: int = 16, chunks=(0, 3, 6, 8), graphs=(1, 8, 32), z_tilings=(2, 1),
                 cache=True):
        """Pick (xchunk, graph_steps, z_tiling) by timing ``trial_steps`` steps of
        each candidate on this batch, then reset every scene."""
        sig = _signature(self.runs)
        path = os.environ.get(_CACHE_ENV)
        table = {}
        if cache and path and os.path.exists(path):
            with open(path) as f:
                table = json.load(f)
        if sig in table and len(table[sig]) == 3:
            self._apply(*table[sig])
            return tuple(table[sig])
        best, best_t = None, float("inf")
        self.autotune_timings = {}  # (xchunk, graph_steps, z_tiling) -> ms per step
        for chunk, zt in [(c, z) for z in z_tilings for c in chunks]:
            self._apply(chunk, 1, zt)
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
                self.autotune_timings[(chunk, g, zt)] = per_step
                if per_step < best_t:
                    best, best_t = (chunk, g, zt), per_step
        self.reset()
        if best is None:
            raise ValueError(f"autotune needs at least {max(trial_steps, 2 * min(graphs))} steps per trial; the "
                             f"batch has {self.total_steps}")
        self._apply(*best)
        if cache and path:
            table[sig] = list(best)
            with open(path, "w") as f:
                json.dump(table, f)
        return best

    def _apply(self, chunk, graph_steps, z_tiling=None):
        z_changed = z_tiling is not None and any(r.program.z_tiling != int(z_tiling) for r in self.runs)
        if z_changed:
            for r in self.runs:
                r.program.z_tiling = int(z_tiling)
        if int(chunk) != self.xchunk or z_changed:
            self.close()
            self.xchunk = int(chunk)
            self._create()
        self.graph_steps = int(graph_steps)


def run_batch_noalloc(runs, steps=None, xchunk: int = 0):
    """Advance every scene of ``runs`` (single-pass ``DeviceRun``s at the same step
    index) by ``steps`` steps through ``mw_batch_run_noalloc``: the batched
    kernels with their tables in the library's static device scratch, so the
    call allocates no device memory.  Returns the runs."""
    runs = list(runs)
    if not runs:
        raise ValueError("a batch needs at least one scene")
    first = runs[0].step_index
    if any(r.step_index != first for r in runs) or any(not r.fused for r in runs):
        raise ValueError("the scenes must be single-pass runs at the same step")
    total = min(r.plan.time_steps for r in runs)
    steps = total - first if steps is None else int(steps)
    lib = runs[0].lib
    progs = (ctypes.POINTER(nat.MwProgram) * len(runs))(*[ctypes.pointer(r.program) for r in runs])
    with torch.cuda.device(runs[0].device):
        rc = lib.mw_batch_run_noalloc(progs, len(runs), int(xchunk), first, steps, nat.current_stream())
    nat.check(rc, lib)
    for r in runs:
        r.step_index = first + steps
    return runs


# ---------------------------------------------------------------- gradients
class BatchGradientSession:
    """Checkpointed forward + reverse sweep of many scenes at once.

    Forward: every scene advances through ``BatchRun`` (one launch per step kernel
    for all scenes), with a snapshot of each scene at every segment start.
    Backward (``reverse="streams"``, the default): each scene's sweep -- restore
    the segment's snapshot, replay it with the design-window capture
    (``mw_capture_steps``), walk it backwards (``mw_reverse_steps``), last segment
    first -- runs on one of ``streams`` HIP streams, each stream's sweeps issued by
    a host thread of its own, so the scenes' sweeps overlap on the GPU and their
    launches are issued side by side; the caller's stream waits for all of them.  ``reverse="kernels"``
    names the batched reverse kernels (``mw_adjoint_batch_*``, all scenes per
    launch), which are switched off: their first device run ended in a GPU
    memory fault whose cause is not established."""

    def __init__(self, plans, device="cuda", windows=None, stride=None, xchunk=0, graph_steps=1, streams=8,
                 reverse="streams"):
        from .adjoint import default_stride, design_window
        if reverse not in ("streams", "kernels"):
            raise ValueError("reverse must be 'streams' or 'kernels'")
        self.plans = list(plans)
        T = {p.time_steps for p in self.plans}
        if len(T) != 1:
            raise ValueError("batched scenes must run the same number of steps")
        self.T = T.pop()
        self.runs = [DeviceRun(p, device, fused=True) for p in self.plans]
        self.batch = BatchRun(self.runs, xchunk=xchunk)
        self.batch.graph_steps = int(graph_steps)
        self.windows = list(windows) if windows is not None else [design_window(p) for p in self.plans]
        self.stride = int(stride or default_stride(self.T))
        self.streams = max(1, int(streams))
        self.reverse = reverse
        self.checkpoints = []
        self.scene_done_s = None

    def outputs(self, si):
        from .adjoint import run_outputs
        return run_outputs(self.runs[si])

    def forward(self):
        from .adjoint import snapshot
        self.batch.reset()
        self.checkpoints = []
        K = self.stride
        for s0 in range(0, self.T, K):
            self.checkpoints.append([snapshot(r) for r in self.runs] if s0 > 0 else None)
            self.batch.run(min(K, self.T - s0))
        return [self.outputs(i) for i in range(len(self.runs))]

    def _adjoint(self, run, win, grads):
        """Adjoint state of one scene seeded with its output gradients scaled to
        unit peak (the sweep is linear in its seeds); returns (adjoint, scale)."""
        from .adjoint import DeviceAdjoint
        grads = list(grads)
        point = (grads.pop(0), grads.pop(0)) if run.point_re is not None else None
        n = len(run.plane_re)
        planes = list(zip(grads[:n], grads[n:2 * n]))
        peak = 0.0
        for t in ([*point] if point else []) + [x for pair in planes for x in pair]:
            peak = max(peak, float(t.detach().abs().max()) if t.numel() else 0.0)
        scale = peak if peak > 0 else 1.0
        sc = lambda t: t.detach().double() / scale
        adj = DeviceAdjoint(run, win, capacity=self.stride)
        adj.zero()
        adj.set_seeds(point=(sc(point[0]), sc(point[1])) if point else None,
                      planes=[(sc(r), sc(i)) for r, i in planes])
        return adj, scale

    def backward(self, grads_per_scene):
        """grads_per_scene[i] follows outputs(i); returns per scene the window
        (g_decay[3], g_curl[3]) as float64 tensors.  ``scene_done_s`` then holds
        each scene's sweep time from the start of the backward pass."""
        if self.reverse == "kernels":
            raise NotImplementedError(
                "the batched reverse kernels are switched off pending an unexplained GPU memory fault; "
                "use reverse='streams' (each scene's sweep on its own stream)")
        from .adjoint import restore
        K, T = self.stride, self.T
        dev = self.runs[0].device
        made = [self._adjoint(run, win, g) for run, win, g in zip(self.runs, self.windows, grads_per_scene)]
        adjs, scales = [a for a, _ in made], [s for _, s in made]
        # the replays run the single-scene step: give them its default tiling (the
        # batch's tuned tiling suits the batched launches), restored afterwards
        tilings = [(r.program.xchunk, r.program.shell_chunk, r.program.z_tiling) for r in self.runs]
        for r in self.runs:
            r.set_tiling(0, 0, 0)
        main = torch.cuda.current_stream(dev)
        pool = [torch.cuda.Stream(device=dev) for _ in range(min(self.streams, len(self.runs)))]
        begin = torch.cuda.Event(enable_timing=True)
        begin.record(main)
        for st in pool:
            st.wait_stream(main)  # snapshots, seeds and zeroed adjoint state are on the caller's stream
        starts = list(range(0, T, K))
        done = [None] * len(self.runs)

        def sweep_stream(w):
            # scenes w, w + len(pool), ... on stream w, issued by this host thread: the
            # sweeps of small scenes are launch-bound, so one thread per stream issues
            # them side by side (the library calls release the interpreter lock)
            st = pool[w]
            with torch.cuda.device(dev), torch.cuda.stream(st):
                for si in range(w, len(self.runs), len(pool)):
                    run, adj = self.runs[si], adjs[si]
                    for ci, s0 in reversed(list(enumerate(starts))):
                        cnt = min(K, T - s0)
                        snaps = self.checkpoints[ci]
                        restore(run, None if snaps is None else snaps[si], s0)
                        adj.capture(s0, cnt)
                        adj.reverse(s0, cnt)
                    ev = torch.cuda.Event(enable_timing=True)
                    ev.record(st)
                    done[si] = ev

        if len(pool) == 1:
            sweep_stream(0)
        else:
            from concurrent.futures import ThreadPoolExecutor
            with ThreadPoolExecutor(max_workers=len(pool)) as ex:
                for f in [ex.submit(sweep_stream, w) for w in range(len(pool))]:
                    f.result()
        for st in pool:
            main.wait_stream(st)
        for run, t in zip(self.runs, tilings):
            run.step_index = T
            run.set_tiling(*t)
        self.adjoints = adjs
        for ev in done:
            ev.synchronize()
        self.scene_done_s = [begin.elapsed_time(ev) * 1e-3 for ev in done]
        return [([g.double() * sc_ for g in a.g_decay], [g.double() * sc_ for g in a.g_curl])
                for a, sc_ in zip(adjs, scales)]


class _BatchGradientFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, session, n_per_scene, *coefs):
        ctx.session = session
        ctx.coef_meta = [(c.device, c.dtype) for c in coefs]
        ctx.n_out = [len(session.outputs(i)) for i in range(len(session.runs))]
        outs = session.forward()
        return tuple(o.detach().clone() for scene in outs for o in scene)

    @staticmethod
    def backward(ctx, *grad_outputs):
        session = ctx.session
        grads, pos = [], 0
        for i, n in enumerate(ctx.n_out):
            outs = session.outputs(i)
            g = [grad_outputs[pos + k] if grad_outputs[pos + k] is not None else torch.zeros_like(outs[k])
                 for k in range(n)]
            grads.append(g)
            pos += n
        res = session.backward(grads)
        flat = []
        for g_decay, g_curl in res:
            flat += list(g_decay) + list(g_curl)
        return (None, None, *[g.to(device=d, dtype=t) for g, (d, t) in zip(flat, ctx.coef_meta)])


def batched_differentiable_run(scenes, plan_kwargs, device="cuda", stride=None, xchunk=0, graph_steps=1, streams=8,
                               reverse="streams"):
    """Differentiable run of many scenes at once.  ``scenes`` is a list of
    ``Scene`` whose material regions carry the densities to differentiate;
    returns (plans, monitors per scene, session)."""
    from ..scene import prepare_scene
    from . import materials as mat
    from .adjoint import design_window, window_coefficients
    from .plan import compile_plan
    from .spectral import assemble_monitor_results
    plans, coefs = [], []
    for scene in scenes:
        dens = [r.density for r in scene.material_regions]
        prepared = prepare_scene(scene)
        eps_node = mat.node_permittivity(prepared, mat.NodeMaterialModel(prepared), dens)
        plan = compile_plan(scene, region_densities=[d.detach() for d in dens], **plan_kwargs)
        lo, hi = design_window(plan)
        c = window_coefficients(plan, eps_node, lo, hi)
        coefs += [d for d, _ in c] + [cu for _, cu in c]
        plans.append(plan)
    session = BatchGradientSession(plans, device, stride=stride, xchunk=xchunk, graph_steps=graph_steps,
                                   streams=streams, reverse=reverse)
    outs = list(_BatchGradientFunction.apply(session, len(scenes), *coefs))
    monitors, pos = [], 0
    for i, (plan, run) in enumerate(zip(plans, session.runs)):
        n = len(session.outputs(i))
        mine = outs[pos:pos + n]
        pos += n
        pr = pi = None
        if run.point_re is not None:
            pr, pi = mine.pop(0), mine.pop(0)
        k = len(run.plane_re)
        qr, qi = mine[:k], mine[k:2 * k]
        if plan.observers is None:
            monitors.append({})
            continue
        if pr is None:
            pr = torch.zeros((len(plan.observers.frequencies), 0), dtype=torch.float32, device=run.device)
            pi = pr
        monitors.append(assemble_monitor_results(plan.observers, pr, pi, qr, qi, plan.obs_window_sum))
    return plans, monitors, session
