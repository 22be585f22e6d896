"""Reverse-mode gradients of monitor outputs with respect to design-region
permittivity.

Pipeline (the reference's ``_FDTDGradientBridge``, fdtd/adjoint/bridge.py:728-853
and :1430-1450; checkpoints fdtd/checkpoint.py:92-125, :365-440, :537-567;
stride fdtd/adjoint/core.py:379-383; segment replay core.py:3495-3565):

1. forward: the device run advances in segments of ``stride`` steps and keeps a
   snapshot of the whole state (E, H, psi, ADE) at every segment start;
2. backward, last segment first: restore the segment's snapshot, replay it with
   ``mw_capture_steps`` (E^n and H^{n+1/2} captured on the design window only,
   not whole states), then ``mw_reverse_steps`` walks the segment backwards,
   seeding the adjoint from the gradients of the observer sums and accumulating
   dL/d(decay_E), dL/d(curl_E) on the window;
3. autograd carries those two window gradients through the float64 coefficient
   formulas, the node->edge averaging and the region density map back to the
   user's density tensors.

The adjoint is the exact transpose of the discrete forward map (the oracle's
reverse step is pinned by a transpose identity and central differences), so
the gradient is that of the discretised problem.  Source patches are held
fixed: a design window may not overlap a source patch, a dispersive pole or
a CPML slab (checked).
"""

from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import materials as mat
from . import native as nat
from .cpml import PSI_NAMES
from .plan import COMPONENTS, FDTDPlan
from .runtime import DeviceRun

E_COMPS = COMPONENTS[:3]


def default_stride(steps: int) -> int:
    """floor(sqrt(T)) steps per checkpoint segment (fdtd/adjoint/core.py:379-383)."""
    return max(1, int(math.isqrt(max(int(steps), 1))))


# ------------------------------------------------------------------ state
def snapshot(run: DeviceRun) -> dict:
    """Clone of the state holding step ``run.step_index``."""
    return {"fields": {c: t.clone() for c, t in run.fields.items()},
            "psi": {name: run.psi_current(name).clone() for name in run.psi},
            "ade": {c: [{k: (v.clone() if v is not None else None) for k, v in s.items()} for s in states]
                    for c, states in run.ade_state.items()},
            "aux": None if run.aux_e is None else (run.aux_e.clone(), run.aux_h.clone())}


def restore(run: DeviceRun, snap: dict | None, step: int) -> None:
    """Load ``snap`` (None = the all-zero initial state) as the state of ``step``
    into the buffers the single-pass step expects for that step's parity."""
    run.step_index = int(step)
    for c, t in run.fields.items():
        t.copy_(snap["fields"][c]) if snap is not None else t.zero_()
    for name in run.psi:
        t = run.psi_current(name)
        t.copy_(snap["psi"][name]) if snap is not None else t.zero_()
    for c, states in run.ade_state.items():
        for si, s in enumerate(states):
            for k, v in s.items():
                if v is not None:
                    v.copy_(snap["ade"][c][si][k]) if snap is not None else v.zero_()
    if run.aux_e is not None:
        if snap is not None:
            run.aux_e.copy_(snap["aux"][0])
            run.aux_h.copy_(snap["aux"][1])
        else:
            run.aux_e.zero_()
            run.aux_h.zero_()


# ------------------------------------------------------------------ window
def design_window(plan: FDTDPlan, margin: int = 1):
    """Node box [lo, hi) whose E coefficients depend on the material regions'
    nodes: the regions' node bounding box widened by ``margin`` (an edge
    coefficient averages its two end nodes)."""
    mask = plan.material_model.region_mask
    if mask is None or not mask.any():
        raise ValueError("the scene has no material-region nodes to differentiate")
    idx = np.nonzero(mask)
    lo = [int(a.min()) - margin for a in idx]
    hi = [int(a.max()) + 1 + margin for a in idx]
    return tuple(lo), tuple(hi)


def check_window(plan: FDTDPlan, lo, hi):
    shape = plan.shape
    for a in range(3):
        if lo[a] < 1 or hi[a] > shape[a] - 1:
            raise ValueError(f"design window axis {'xyz'[a]} [{lo[a]}, {hi[a]}) touches the grid faces")
    if plan.uses_cpml:
        for name in PSI_NAMES[:6]:
            axis, low, hs, hl = plan.psi_layout[name]
            if lo[axis] < low or (hl > 0 and hi[axis] > hs):
                raise ValueError(f"design window [{lo}, {hi}) overlaps the CPML layers along {'xyz'[axis]}")
    box = tuple(slice(l, h) for l, h in zip(lo, hi))
    for c in E_COMPS:
        for pole in plan.ade[c]["poles"]:
            if np.any(np.asarray(pole["drive"])[box] != 0):
                raise NotImplementedError("gradients through dispersive poles inside a design window")
    for t in plan.point_terms + plan.e_terms + plan.h_terms + plan.line_e_terms + plan.line_h_terms:
        off, shp = t["offset"], t["patch"].shape
        if all(off[a] < hi[a] + 1 and off[a] + shp[a] > lo[a] - 1 for a in range(3)):
            raise NotImplementedError("a source patch overlaps the design window (source patches are held fixed)")


def window_coefficients(plan: FDTDPlan, eps_node: torch.Tensor, lo, hi):
    """(decay, curl) of each E component on the window box, float64 torch,
    differentiable in ``eps_node`` (the float64 formulas of the forward plan)."""
    out = []
    sigma = plan.material_model.sigma_e
    for ci, c in enumerate(E_COMPS):
        sl = [slice(l, h) for l, h in zip(lo, hi)]
        sl[ci] = slice(lo[ci], hi[ci] + 1)
        node = eps_node[tuple(sl)]
        eps = mat.average_node_to_component(node, c)
        sig = mat.average_node_to_component(sigma[tuple(sl)], c)
        out.append(mat.electric_coefficients(eps, sig, plan.dt))
    return out


# ------------------------------------------------------------------ adjoint
class DeviceAdjoint:
    """Adjoint state of one ``DeviceRun`` plus the design-window capture."""

    def __init__(self, run: DeviceRun, window=None, capacity: int = 1):
        self.run = run
        plan, dev = run.plan, run.device
        self.lib = run.lib
        self.fields = {c: torch.zeros(plan.field_shape(c), dtype=torch.float32, device=dev) for c in COMPONENTS}
        self.psi = {name: torch.zeros_like(t) for name, t in run.psi.items()}
        self.ade = {c: [{k: (torch.zeros_like(v) if v is not None else None) for k, v in s.items()} for s in states]
                    for c, states in run.ade_state.items()}
        adj = nat.MwAdjoint()
        for a in range(3):
            adj.fields.e[a] = nat.ptr(self.fields[COMPONENTS[a]])
            adj.fields.h[a] = nat.ptr(self.fields[COMPONENTS[a + 3]])
        for q, name in enumerate(PSI_NAMES):
            if name in self.psi and self.psi[name].numel():
                adj.psi[q] = nat.ptr(self.psi[name])
        for ci, c in enumerate(E_COMPS):
            for pi, s in enumerate(self.ade.get(c, [])):
                adj.current[ci][pi] = nat.ptr(s["current"])
                adj.polarization[ci][pi] = nat.ptr(s["polarization"]) if s["polarization"] is not None else None
        self.adjoint = adj
        self._seed_refs = []
        self.window = None
        if window is not None:
            lo, hi = (tuple(int(v) for v in window[0]), tuple(int(v) for v in window[1]))
            check_window(plan, lo, hi)
            wshape = tuple(h - l for l, h in zip(lo, hi))
            w_cells = int(np.prod(wshape))
            wh_cells = int(np.prod([s + 1 for s in wshape]))
            capacity = int(capacity)
            self.e_store = torch.empty(((capacity + 1) * 3 * w_cells,), dtype=torch.float32, device=dev)
            self.h_store = torch.empty((capacity * 3 * wh_cells,), dtype=torch.float32, device=dev)
            self.g_decay = [torch.zeros(wshape, dtype=torch.float32, device=dev) for _ in range(3)]
            self.g_curl = [torch.zeros(wshape, dtype=torch.float32, device=dev) for _ in range(3)]
            w = nat.MwWindow()
            for a in range(3):
                w.lo[a] = lo[a]
                w.hi[a] = hi[a]
                w.g_decay[a] = nat.ptr(self.g_decay[a])
                w.g_curl[a] = nat.ptr(self.g_curl[a])
            w.capacity = capacity
            w.e_store = nat.ptr(self.e_store)
            w.h_store = nat.ptr(self.h_store)
            self.window = w
            self.window_box = (lo, hi)

    def zero(self):
        for t in list(self.fields.values()) + list(self.psi.values()):
            t.zero_()
        for states in self.ade.values():
            for s in states:
                for v in s.values():
                    if v is not None:
                        v.zero_()
        if self.window is not None:
            for t in self.g_decay + self.g_curl:
                t.zero_()

    def set_seeds(self, point=None, planes=None, dft=None):
        """Gradients of the observer sums: ``point`` = (re, im) of shape (F, n_points),
        ``planes`` = list of (re, im) per plane (None to skip), ``dft`` = {comp: (re, im)}."""
        adj = self.adjoint
        refs = []

        def p(t):
            t = t.to(device=self.run.device, dtype=torch.float32).contiguous()
            refs.append(t)
            return nat.ptr(t)

        adj.g_point_re = adj.g_point_im = None
        if point is not None:
            adj.g_point_re, adj.g_point_im = p(point[0]), p(point[1])
        for q in range(nat.MW_MAX_PLANES):
            adj.g_plane_re[q] = adj.g_plane_im[q] = None
        for q, pair in enumerate(planes or []):
            if pair is not None:
                adj.g_plane_re[q], adj.g_plane_im[q] = p(pair[0]), p(pair[1])
        for a in range(3):
            adj.g_dft_re[a] = adj.g_dft_im[a] = None
        for c, (re, im) in (dft or {}).items():
            a = E_COMPS.index(c)
            adj.g_dft_re[a], adj.g_dft_im[a] = p(re), p(im)
        self._seed_refs = refs

    def capture(self, first: int, count: int):
        with torch.cuda.device(self.run.device):
            rc = self.lib.mw_capture_steps(ctypes.byref(self.run.program), int(first), int(count),
                                           ctypes.byref(self.window), nat.current_stream())
        nat.check(rc, self.lib)

    def reverse(self, first: int, count: int, gradients: bool = True):
        win = ctypes.byref(self.window) if (gradients and self.window is not None) else None
        with torch.cuda.device(self.run.device):
            rc = self.lib.mw_reverse_steps(ctypes.byref(self.run.program), ctypes.byref(self.adjoint), win,
                                           int(first), int(count), nat.current_stream())
        nat.check(rc, self.lib)


def run_outputs(run: DeviceRun):
    """The observer sums of a run, in the order the autograd boundary returns them:
    point re/im, plane re (each), plane im (each), full-field DFT re, im."""
    out = []
    if run.point_re is not None:
        out += [run.point_re, run.point_im]
    out += list(run.plane_re) + list(run.plane_im)
    out += [run.dft_re[c] for c in E_COMPS if c in run.dft_re] + [run.dft_im[c] for c in E_COMPS if c in run.dft_im]
    return out


class GradientSession:
    """Checkpointed forward + reverse sweep for one plan with a design window."""

    def __init__(self, plan: FDTDPlan, device="cuda", window=None, stride=None, fused=None, tune=False):
        self.plan = plan
        self.run = DeviceRun(plan, device, fused=fused)
        if tune and self.run.fused and plan.time_steps >= 200:
            # a few trial steps of each candidate tiling (the run is reset afterwards)
            self.run.autotune(trial_steps=8, xchunks=(0, 6, 8), shell_chunks=(2, 4), z_tilings=(1, 2, 3, 4))
        self.window = window if window is not None else design_window(plan)
        self.stride = int(stride or default_stride(plan.time_steps))
        self.checkpoints = {}

    def outputs(self):
        return run_outputs(self.run)

    def forward(self):
        run, T, K = self.run, self.plan.time_steps, self.stride
        run.reset()
        self.checkpoints = {}
        for s0 in range(0, T, K):
            if s0 > 0:
                self.checkpoints[s0] = snapshot(run)
            run.run(min(K, T - s0))
        return self.outputs()

    def backward(self, grad_outputs):
        """grad_outputs follow ``outputs()``; returns (g_decay[3], g_curl[3]) on the
        window as float64 tensors."""
        run, T, K = self.run, self.plan.time_steps, self.stride
        grads = list(grad_outputs)
        point = None
        if run.point_re is not None:
            point = (grads.pop(0), grads.pop(0))
        n_planes = len(run.plane_re)
        pre, pim = grads[:n_planes], grads[n_planes:2 * n_planes]
        grads = grads[2 * n_planes:]
        planes = [(r, i) for r, i in zip(pre, pim)]
        dft = {}
        dcomps = [c for c in E_COMPS if c in run.dft_re]
        for c, r, i in zip(dcomps, grads[:len(dcomps)], grads[len(dcomps):]):
            dft[c] = (r, i)
        # the sweep is linear in its seeds: run it on seeds scaled to unit peak so
        # float32 adjoint fields and window sums stay far from overflow whatever
        # the loss's magnitude, and rescale the window gradients in float64
        peak = 0.0
        for t in ([*point] if point else []) + [x for pair in planes for x in pair] + \
                [x for pair in dft.values() for x in pair]:
            peak = max(peak, float(t.detach().abs().max()) if t.numel() else 0.0)
        scale = peak if peak > 0 else 1.0
        sc = lambda t: (t.detach().double() / scale)
        point = (sc(point[0]), sc(point[1])) if point else None
        planes = [(sc(r), sc(i)) for r, i in planes]
        dft = {c: (sc(r), sc(i)) for c, (r, i) in dft.items()}
        adj = DeviceAdjoint(run, self.window, capacity=K)
        adj.zero()
        adj.set_seeds(point=point, planes=planes, dft=dft)
        starts = list(range(0, T, K))
        for s0 in reversed(starts):
            count = min(K, T - s0)
            restore(run, self.checkpoints.get(s0), s0)
            adj.capture(s0, count)
            adj.reverse(s0, count)
        run.step_index = T
        self.adjoint = adj
        return [g.double() * scale for g in adj.g_decay], [g.double() * scale for g in adj.g_curl]


def _zeros_for(t):
    return torch.zeros_like(t)


class _FDTDGradientFunction(torch.autograd.Function):
    """Autograd boundary (fdtd/adjoint/bridge.py:1430-1450): inputs are the window
    coefficients (decay_E x3, curl_E x3, float64, carrying the graph back to the
    densities), outputs the observer sums of the device run."""

    @staticmethod
    def forward(ctx, session, *coefs):
        ctx.session = session
        ctx.coef_meta = [(c.device, c.dtype) for c in coefs]
        outs = session.forward()
        return tuple(o.detach().clone() for o in outs)

    @staticmethod
    def backward(ctx, *grad_outputs):
        session = ctx.session
        outs = session.outputs()
        grads = [g if g is not None else _zeros_for(o) for g, o in zip(grad_outputs, outs)]
        g_decay, g_curl = session.backward(grads)
        # each gradient lives where its coefficient input does (the host float64 graph)
        return (None, *[g.to(device=d, dtype=t) for g, (d, t) in zip(list(g_decay) + list(g_curl), ctx.coef_meta)])


def differentiable_run(scene, region_densities=None, device="cuda", stride=None, fused=None, tune=False,
                       **plan_kwargs):
    """Run ``scene`` on the device with gradients enabled for its material
    regions.  Returns (plan, monitors) where the monitor payloads are torch
    tensors differentiable in the regions' densities.  ``tune``: time a few
    single-pass tilings on the run first (the choice is kept for later runs of
    the same signature in this process)."""
    from .plan import compile_plan
    from .spectral import assemble_monitor_results
    from ..scene import prepare_scene
    regions = list(scene.material_regions)
    if region_densities is None:
        region_densities = [r.density for r in regions]
    prepared = prepare_scene(scene)
    model = mat.NodeMaterialModel(prepared)
    eps_node = mat.node_permittivity(prepared, model, region_densities)
    plan = compile_plan(scene, region_densities=[d.detach() for d in region_densities], **plan_kwargs)
    window = design_window(plan)
    session = GradientSession(plan, device, window=window, stride=stride, fused=fused, tune=tune)
    lo, hi = window
    coefs = window_coefficients(plan, eps_node, lo, hi)
    # the device runs the plan's float32 coefficients; they must be the values
    # the differentiable formulas give on the window
    for ci, c in enumerate(E_COMPS):
        box = tuple(slice(l, h) for l, h in zip(lo, hi))
        dev_curl = np.broadcast_to(plan.coef[c][1], plan.field_shape(c))[box]
        if not np.allclose(coefs[ci][1].detach().cpu().numpy(), dev_curl, rtol=1e-12, atol=0):
            raise RuntimeError("window coefficients diverge from the compiled plan")
    flat = [d for d, _ in coefs] + [cu for _, cu in coefs]
    outs = _FDTDGradientFunction.apply(session, *flat)
    outs = list(outs)
    run = session.run
    point_re = point_im = None
    if run.point_re is not None:
        point_re, point_im = outs.pop(0), outs.pop(0)
    n_planes = len(run.plane_re)
    plane_re, plane_im = outs[:n_planes], outs[n_planes:2 * n_planes]
    monitors = {}
    if plan.observers is not None:
        nf = len(plan.observers.frequencies)
        if point_re is None:
            point_re = torch.zeros((nf, 0), dtype=torch.float32, device=session.run.device)
            point_im = point_re
        monitors = assemble_monitor_results(plan.observers, point_re, point_im, plane_re, plane_im,
                                            plan.obs_window_sum)
    return plan, monitors, session
