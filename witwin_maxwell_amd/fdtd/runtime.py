"""Device runtime: uploads an ``FDTDPlan`` into HBM and drives the native step loop.

Replaces the reference's per-kernel Python orchestration
(fdtd/runtime/stepping.py:2153-2416) with one native call per chunk of steps
(``mw_run_steps``); per-step host scalars (source signals, DFT/observer weights)
become float32 step tables in device memory, so a chunk carries no host input.
Torch owns every buffer; the kernels run on torch's current stream.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import native as nat
from .cpml import PSI_NAMES, PSI_SPECS, compact_shape
from .materials import pole_activity, pole_weight_codes
from .plan import COMPONENTS, FDTDPlan

_KIND = {"debye": 0, "drude": 1, "lorentz": 2}

# tilings chosen by DeviceRun.autotune in this process, per run signature
_TUNED: dict = {}


def _f32(a, device):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device=device)


def _uniform_value(arr):
    """The float32 value of a coefficient array that holds one value, else None."""
    a = np.asarray(arr)
    if a.size and all(s == 0 for s in a.strides):  # a broadcast scalar
        return float(np.float32(a.flat[0]))
    a = a.astype(np.float32)
    if a.size and a.min() == a.max():
        return float(a.flat[0])
    return None


class DeviceRun:
    """One scene resident in HBM.

    ``fused=None`` (or True) runs the single-pass step (E, H and the H-side psi
    double-buffered, one set of launches per step); ``fused=False`` the two-kernel
    in-place path.  Both handle every face mode, periodic included.

    ``plane_stage_depth`` (0 = off): the samples of z-normal plane monitors are
    copied into a stage of that many steps and added to their sums once per block
    of steps, instead of the sums being read and written after every step."""

    def __init__(self, plan: FDTDPlan, device="cuda", fused=None, plane_stage_depth=32):
        self.plan = plan
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise RuntimeError("the FDTD runtime runs on a HIP device; got " + str(self.device))
        self.lib = nat.load_library()
        dev = self.device
        self.keep = []
        nx, ny, nz = plan.shape
        self.fused = True if fused is None else bool(fused)
        self._fields = [{c: torch.zeros(plan.field_shape(c), dtype=torch.float32, device=dev) for c in COMPONENTS}]
        if self.fused:
            self._fields.append({c: torch.zeros_like(t) for c, t in self._fields[0].items()})

        prog = nat.MwProgram()
        g = prog.grid
        g.nx, g.ny, g.nz = nx, ny, nz
        for a in range(3):
            g.bc_lo[a] = plan.bc_lo[a]
            g.bc_hi[a] = plan.bc_hi[a]
            g.inv_e[a] = self._dev(plan.inv_e[a])
            g.inv_h[a] = self._dev(plan.inv_h[a])
        cf = prog.coeffs
        self.uniform = {}
        self.coef_tensors = {}  # (component, "decay"|"curl") -> device array, where not uniform
        for ci, c in enumerate(COMPONENTS):
            dec, curl = plan.coef[c]
            for name, arr, tgt_arr, tgt_val in (("decay", dec, cf.decay, cf.decay_value),
                                                ("curl", curl, cf.curl, cf.curl_value)):
                u = _uniform_value(arr)
                self.uniform[(c, name)] = u
                if u is not None:
                    tgt_arr[ci] = None
                    tgt_val[ci] = u
                else:
                    tgt_arr[ci] = self._dev(np.broadcast_to(arr, plan.field_shape(c)))
                    self.coef_tensors[(c, name)] = self.keep[-1]
                    tgt_val[ci] = 0.0
        cp = prog.cpml
        self.psi = {}
        if plan.uses_cpml:
            cp.enabled = 1
            for a, ax in enumerate("xyz"):
                p = plan.cpml[ax]
                cp.inv_kappa_e[a] = self._dev(1.0 / p["kappa_e"])
                cp.b_e[a] = self._dev(p["b_e"])
                cp.c_e[a] = self._dev(p["c_e"])
                cp.inv_kappa_h[a] = self._dev(1.0 / p["kappa_h"])
                cp.b_h[a] = self._dev(p["b_h"])
                cp.c_h[a] = self._dev(p["c_h"])
            self._psi_alt = {}
            for q, name in enumerate(PSI_NAMES):
                comp, _ = PSI_SPECS[name]
                lay = plan.psi_layout[name]
                shape = compact_shape(plan.field_shape(comp), lay)
                t = torch.zeros(shape, dtype=torch.float32, device=dev)
                self.psi[name] = t
                cp.psi[q] = nat.ptr(t) if t.numel() else None
                cp.psi_low[q] = lay[1]
                cp.psi_high_start[q] = lay[2]
                cp.psi_high_len[q] = lay[3]
                if self.fused and q >= 6:
                    alt = torch.zeros_like(t)
                    self._psi_alt[name] = alt
                    prog.psi_h_alt[q - 6] = nat.ptr(alt) if alt.numel() else None
        ad = prog.ade
        ad.dt = plan.dt
        self.ade_state = {}
        has_ade = False
        for ci, c in enumerate(("Ex", "Ey", "Ez")):
            poles = plan.ade[c]["poles"]
            if len(poles) > nat.MW_MAX_POLES:
                raise ValueError(f"at most {nat.MW_MAX_POLES} poles per component")
            ad.count[ci] = len(poles)
            if poles:
                has_ade = True
                ad.inv_eps[ci] = self._dev(plan.ade[c]["inv_eps"])
                ad.active[ci] = self._dev(pole_activity(poles), dtype=np.uint8)
                codes = pole_weight_codes(poles)
                if codes is not None:
                    ad.weight[ci] = self._dev(codes, dtype=np.uint8)
            states = []
            for pi, pole in enumerate(poles):
                pl = ad.poles[ci][pi]
                pl.kind = _KIND[pole["kind"]]
                pl.decay = pole["decay"]
                pl.restoring = pole["restoring"]
                pl.drive_scale = pole.get("drive_scale", 0.0)
                pl.drive = self._dev(pole["drive"])
                cur = torch.zeros(plan.field_shape(c), dtype=torch.float32, device=dev)
                pol = torch.zeros_like(cur) if pole["kind"] != "drude" else None
                pl.current = nat.ptr(cur)
                pl.polarization = nat.ptr(pol) if pol is not None else None
                states.append({"current": cur, "polarization": pol})
            self.ade_state[c] = states
        prog.has_ade = 1 if has_ade else 0
        for f, fd in ((prog.fields, self._fields[0]), (prog.fields_alt, self._fields[-1])):
            for a in range(3):
                f.e[a] = nat.ptr(fd[COMPONENTS[a]])
                f.h[a] = nat.ptr(fd[COMPONENTS[a + 3]])
        prog.fused = 1 if self.fused else 0
        prog.dt = plan.dt
        prog.total_steps = plan.time_steps

        # ---- TFSF auxiliary line
        self.aux_e = self.aux_h = None
        if plan.aux is not None:
            aux = plan.aux
            n_e = int(aux["n_e"])
            self.aux_e = torch.zeros(n_e, dtype=torch.float32, device=dev)
            self.aux_h = torch.zeros(n_e - 1, dtype=torch.float32, device=dev)
            line = prog.aux
            line.n_e = n_e
            line.e = nat.ptr(self.aux_e)
            line.h = nat.ptr(self.aux_h)
            line.decay_e = self._dev(aux["decay_e"])
            line.curl_e = self._dev(aux["curl_e"])
            line.decay_h = self._dev(aux["decay_h"])
            line.curl_h = self._dev(aux["curl_h"])
            line.source = self._dev(aux["source"])

        # ---- source terms
        cidx = {c: i for i, c in enumerate(COMPONENTS)}
        e_terms, h_terms = [], []
        for t in plan.line_h_terms:
            h_terms.append(self._term_line(cidx[t["field"]], t, self.aux_e))
        for t in plan.line_e_terms:
            e_terms.append(self._term_line(cidx[t["field"]], t, self.aux_h))
        for ti, t in enumerate(plan.point_terms):
            e_terms.append(self._term_additive(cidx[t["field"]], t, plan.signals[ti]))
        for t in plan.e_terms:
            e_terms.append(self._term_shifted(cidx[t["field"]], t, 0.5 * plan.dt))
        for t in plan.h_terms:
            h_terms.append(self._term_shifted(cidx[t["field"]], t, 0.0))
        if len(e_terms) > nat.MW_MAX_TERMS or len(h_terms) > nat.MW_MAX_TERMS:
            raise ValueError(f"at most {nat.MW_MAX_TERMS} source terms per field family")
        self._e_terms = (nat.MwTerm * max(1, len(e_terms)))(*e_terms)
        self._h_terms = (nat.MwTerm * max(1, len(h_terms)))(*h_terms)
        prog.n_e_terms = len(e_terms)
        prog.n_h_terms = len(h_terms)
        prog.e_terms = ctypes.cast(self._e_terms, ctypes.POINTER(nat.MwTerm))
        prog.h_terms = ctypes.cast(self._h_terms, ctypes.POINTER(nat.MwTerm))
        # device copies of the descriptors for the single-pass kernel
        for arr, attr in ((self._e_terms, "e_terms_dev"), (self._h_terms, "h_terms_dev")):
            raw = torch.frombuffer(bytearray(ctypes.string_at(ctypes.addressof(arr), ctypes.sizeof(arr))),
                                   dtype=torch.uint8).to(dev)
            self.keep.append(raw)
            setattr(prog, attr, nat.ptr(raw))

        # ---- observers
        obs = plan.observers
        self.point_re = self.point_im = None
        self.plane_re, self.plane_im = [], []
        planes = []
        if obs is not None:
            nf = len(obs.frequencies)
            prog.nf_obs = nf
            ec, es, hc, hs = plan.obs_weights
            prog.obs_e_cos = self._dev(ec)
            prog.obs_e_sin = self._dev(es)
            prog.obs_h_cos = self._dev(hc)
            prog.obs_h_sin = self._dev(hs)
            shapes = {c: plan.field_shape(c) for c in COMPONENTS}
            for (comp, axis, index) in obs.planes:
                pshape = obs.plane_shape(shapes, (comp, axis, index))
                re = torch.zeros((nf,) + pshape, dtype=torch.float32, device=dev)
                im = torch.zeros_like(re)
                self.plane_re.append(re)
                self.plane_im.append(im)
                po = nat.MwPlaneObs()
                po.component = cidx[comp]
                po.axis = "xyz".index(axis)
                po.index = index
                po.re = nat.ptr(re)
                po.im = nat.ptr(im)
                planes.append(po)
            if len(planes) > nat.MW_MAX_PLANES:
                raise ValueError(f"at most {nat.MW_MAX_PLANES} observer planes")
            if obs.points:
                comps = np.array([cidx[c] for c, _ in obs.points], dtype=np.int32)
                lin = np.array([np.ravel_multi_index(idx, plan.field_shape(c)) for c, idx in obs.points],
                               dtype=np.int32)
                self.point_comp_t = torch.as_tensor(comps, device=dev)
                self.point_lin_t = torch.as_tensor(lin, device=dev)
                self.point_re = torch.zeros((nf, len(obs.points)), dtype=torch.float32, device=dev)
                self.point_im = torch.zeros_like(self.point_re)
                prog.point_component = nat.ptr(self.point_comp_t)
                prog.point_linear = nat.ptr(self.point_lin_t)
                prog.point_re = nat.ptr(self.point_re)
                prog.point_im = nat.ptr(self.point_im)
                prog.n_points = len(obs.points)
        self._planes = (nat.MwPlaneObs * max(1, len(planes)))(*planes)
        prog.planes = ctypes.cast(self._planes, ctypes.POINTER(nat.MwPlaneObs))
        prog.n_planes = len(planes)
        # z-normal planes: a row per step of each distinct (component, index)
        self.plane_stage = None
        z_pairs = {}
        for q, po in enumerate(planes):
            if po.axis == 2:
                z_pairs.setdefault((po.component, po.index), int(self.plane_re[q][0].numel()))
        z_cells = sum(z_pairs.values())
        depth = int(plane_stage_depth)
        if z_cells > 0 and depth > 0:
            if depth > 64:
                raise ValueError("plane_stage_depth must be at most 64")
            self.plane_stage = torch.zeros((depth, z_cells), dtype=torch.float32, device=dev)
            prog.plane_stage = nat.ptr(self.plane_stage)
            prog.plane_stage_floats = self.plane_stage.numel()
            prog.plane_stage_depth = depth

        # ---- full-field DFT
        self.dft_re, self.dft_im = {}, {}
        if plan.dft_frequencies:
            nf = len(plan.dft_frequencies)
            prog.nf_dft = nf
            prog.dft_cos = self._dev(plan.dft_weights[0])
            prog.dft_sin = self._dev(plan.dft_weights[1])
            for a, c in enumerate(("Ex", "Ey", "Ez")):
                re = torch.zeros((nf,) + plan.field_shape(c), dtype=torch.float32, device=dev)
                im = torch.zeros_like(re)
                self.dft_re[c] = re
                self.dft_im[c] = im
                prog.dft_re[a] = nat.ptr(re)
                prog.dft_im[a] = nat.ptr(im)
        self.program = prog
        self.step_index = 0

    @property
    def fields(self):
        """Current field tensors (the double buffer holding step ``step_index``)."""
        return self._fields[self.step_index & 1] if self.fused else self._fields[0]

    def psi_current(self, name):
        if self.fused and name.startswith("psi_h") and (self.step_index & 1):
            return self._psi_alt[name]
        return self.psi[name]

    # ------------------------------------------------------------ helpers
    def _dev(self, arr, dtype=np.float32):
        t = torch.as_tensor(np.ascontiguousarray(arr, dtype=dtype), device=self.device)
        self.keep.append(t)
        return nat.ptr(t)

    def _term_common(self, comp, t):
        term = nat.MwTerm()
        term.component = comp
        for a in range(3):
            term.offset[a] = int(t["offset"][a])
            term.shape[a] = int(t["patch"].shape[a])
        term.patch = self._dev(t["patch"])
        return term

    def _term_additive(self, comp, t, signal):
        term = self._term_common(comp, t)
        term.kind = 0
        term.signal_table = self._dev(signal)
        return term

    def _term_line(self, comp, t, line):
        term = self._term_common(comp, t)
        term.kind = 2
        term.line = nat.ptr(line)
        term.line_axis = int(t["line_axis"])
        term.line_base = int(t["line_base"])
        term.line_step = int(t["line_step"])
        term.line_len = int(line.numel())
        return term

    def _term_shifted(self, comp, t, t_offset):
        term = self._term_common(comp, t)
        term.kind = 1
        term.delay = self._dev(t["delay"], dtype=np.float64)
        vals = torch.zeros(t["patch"].shape, dtype=torch.float32, device=self.device)
        self.keep.append(vals)
        term.values = nat.ptr(vals)
        st = t["st"]
        term.wave_kind = st["kind_code"]
        term.t_offset = t_offset
        term.frequency = st["frequency"]
        term.fwidth = st["fwidth"]
        term.amplitude = st["amplitude"]
        term.phase = st["phase"]
        term.delay0 = st["delay"]
        return term

    # ------------------------------------------------------------ stepping
    def run(self, count=None):
        count = self.plan.time_steps - self.step_index if count is None else int(count)
        if count <= 0:
            return self
        with torch.cuda.device(self.device):
            rc = self.lib.mw_run_steps(ctypes.byref(self.program), self.step_index, count, nat.current_stream())
        nat.check(rc, self.lib)
        self.step_index += count
        return self

    def set_tiling(self, xchunk: int = 0, shell_chunk: int = 0, z_tiling: int = 0):
        """Single-pass tiling: x planes per block of the interior and PML-shell
        launches (0 = the library's choice) and how the z ends are tiled (0 or 2:
        interior tiles aligned to the z PML slabs, z ends in the per-cell shell;
        1: z-edge tiles beside the interior tiles; 3: aligned interior tiles and
        both z ends of two rows per wavefront, where they fit, else as 2; 4: as 3
        with the z-end launch on a second stream beside the interior launch)."""
        self.program.xchunk = int(xchunk)
        self.program.shell_chunk = int(shell_chunk)
        self.program.z_tiling = int(z_tiling)
        return self

    def autotune(self, trial_steps: int = 12, xchunks=(0, 3, 4, 6, 8), shell_chunks=(2, 4), z_tilings=(1, 2, 3, 4),
                 cache=True, prefer=(), margin=0.01):
        """Pick the single-pass tiling by timing ``trial_steps`` steps of each
        candidate on this run, then reset the run to step 0.  Among candidates
        within ``margin`` of the fastest, the first of ``prefer`` (tilings the
        caller has other data for) is taken.  The choice is cached per (shape,
        CPML, dispersive, DFT) signature for the rest of the process and in the
        JSON file named by WITWIN_MAXWELL_AMD_TUNE_CACHE, when set.  Returns
        (xchunk, shell_chunk, z_tiling)."""
        import json
        import os
        if not self.fused:
            return (0, 0, 0)
        p = self.plan
        sig = f"single|{p.shape}|{int(p.uses_cpml)}|{int(p.dispersive)}|{len(p.dft_frequencies)}"
        path = os.environ.get("WITWIN_MAXWELL_AMD_TUNE_CACHE")
        table = {}
        if cache and path and os.path.exists(path):
            with open(path) as f:
                table = json.load(f)
        if cache and sig in _TUNED:
            self.set_tiling(*_TUNED[sig])
            return tuple(_TUNED[sig])
        if sig in table:
            _TUNED[sig] = tuple(table[sig])
            self.set_tiling(*table[sig])
            return tuple(table[sig])
        steps = min(int(trial_steps), max(1, p.time_steps - 1))
        timed = {}
        cands = [(xc, sc, zt) for zt in z_tilings for xc in xchunks for sc in shell_chunks]
        cands += [tuple(int(v) for v in c) for c in prefer if tuple(int(v) for v in c) not in cands]
        for cand in cands:
            self.set_tiling(*cand)
            self.reset()
            self.run(1)  # warm-up
            start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            start.record()
            self.run(min(steps, p.time_steps - self.step_index))
            end.record()
            end.synchronize()
            timed[cand] = start.elapsed_time(end)
        # ms per step of every candidate, for the caller's records
        self.autotune_timings = {c: t / steps for c, t in timed.items()}
        best = min(timed, key=timed.get)
        for c in prefer:
            c = tuple(int(v) for v in c)
            if timed.get(c, float("inf")) <= timed[best] * (1.0 + margin):
                best = c
                break
        self.set_tiling(*best)
        self.reset()
        if cache:
            _TUNED[sig] = tuple(best)
        if cache and path:
            table[sig] = list(best)
            with open(path, "w") as f:
                json.dump(table, f)
        return best

    def set_coefficients(self, component, name, box, values):
        """Overwrite the device ``name`` ("decay"/"curl") coefficients of
        ``component`` on the index box ``box`` (a tuple of slices) with ``values``
        (e.g. a perturbed design window).  The component's coefficient must be
        stored as an array (not uniform)."""
        t = self.coef_tensors.get((component, name))
        if t is None:
            raise ValueError(f"{component} {name} is uniform on this run: no array to overwrite")
        t[box] = torch.as_tensor(np.asarray(values), dtype=torch.float32, device=self.device)

    def reset(self):
        for fd in self._fields:
            for t in fd.values():
                t.zero_()
        for t in list(self.psi.values()) + list(getattr(self, "_psi_alt", {}).values()):
            t.zero_()
        for states in self.ade_state.values():
            for s in states:
                s["current"].zero_()
                if s["polarization"] is not None:
                    s["polarization"].zero_()
        for t in self.plane_re + self.plane_im + list(self.dft_re.values()) + list(self.dft_im.values()):
            t.zero_()
        if self.plane_stage is not None:
            self.plane_stage.zero_()
        if self.aux_e is not None:
            self.aux_e.zero_()
            self.aux_h.zero_()
        if self.point_re is not None:
            self.point_re.zero_()
            self.point_im.zero_()
        self.step_index = 0

    # ------------------------------------------------------------ outputs
    def monitor_results(self):
        from .spectral import assemble_monitor_results
        obs = self.plan.observers
        if obs is None:
            return {}
        pr = self.point_re.double().cpu().numpy() if self.point_re is not None else np.zeros((len(obs.frequencies), 0))
        pi = self.point_im.double().cpu().numpy() if self.point_im is not None else np.zeros((len(obs.frequencies), 0))
        return assemble_monitor_results(obs, pr, pi, [t.double().cpu().numpy() for t in self.plane_re],
                                        [t.double().cpu().numpy() for t in self.plane_im], self.plan.obs_window_sum)

    def psi_dense(self, name):
        """Expand a slab psi array to the field shape (psi is zero outside the slabs)."""
        comp, axis = PSI_SPECS[name]
        _, low, hs, hl = self.plan.psi_layout[name]
        dense = torch.zeros(self.plan.field_shape(comp), dtype=torch.float32, device=self.device)
        t = self.psi_current(name)
        if low:
            dense.narrow(axis, 0, low).copy_(t.narrow(axis, 0, low))
        if hl:
            dense.narrow(axis, hs, hl).copy_(t.narrow(axis, low, hl))
        return dense
