"""Monitor sampling and payload assembly of the oracle -- TEST INFRASTRUCTURE ONLY.

An independent float64 restatement of how the reference turns field samples
into monitor payloads, so that the parity tests compare the product's monitor
code with this one rather than with itself.  Nothing here imports the product's
spectral module; the oracle decides which samples it records, computes its own
DFT weights and assembles its own payloads and fluxes.

Reference behaviour restated:
  * Yee sample coordinates per component ........ fdtd/coords.py:9-27
  * spectral start step / window weights ......... fdtd/runtime/spectral.py:227-297
  * DFT phase (+ sign, step n at t = n dt), H retarded by omega dt / 2
                                                   fdtd/runtime/spectral.py:197-220, :558-611
  * point monitors: nearest sample per axis ...... fdtd/observers.py:275-282
  * plane monitors: the two straddling planes, linear weights; flux planes
    snapped to the nearest node plane ............ fdtd/observers.py:296-405
  * payload scale 2 / sum(w) ..................... fdtd/observers.py:236-241
  * alignment onto cell centres, crop to the physical domain
                                                   fdtd/observers.py:25-37, :126-200
  * flux = sum_cells 1/2 Re((E x H*) . n) * exact primal cell widths
                                                   fdtd/quadrature.py:78-110, :112-174
Pinned by tests/test_monitor_pinning.py (the flux of a closed-form tangential
field equals its analytic integral; point and interpolated plane payloads equal
direct evaluations) and tests/test_gpu_power.py (a soft plane wave's flux is its
analytic incident power within 2%).
Only tests/ import this module.
"""

from __future__ import annotations

import math

import numpy as np

C0 = 299792458.0
_AX = {"x": 0, "y": 1, "z": 2}
_PLANE_AXES = {"x": ("y", "z"), "y": ("x", "z"), "z": ("x", "y")}


def yee_coords(nodes, comp):
    """(x, y, z) sample coordinates of component ``comp``: an E component sits at
    cell centres along its own axis and on nodes across it; an H component the
    other way round."""
    half = {a: 0.5 * (nodes[a][1:] + nodes[a][:-1]) for a in "xyz"}
    own = comp[1]
    out = []
    for a in "xyz":
        on_half = (a == own) if comp[0] == "E" else (a != own)
        out.append(half[a] if on_half else nodes[a])
    return tuple(out)


def start_step(frequency, dt, window, pulsed, physical_range):
    """First accumulated step: 0 for pulses or no window, otherwise after the
    longer of 15 periods and 5 round trips across the largest physical extent."""
    if pulsed or window == "none":
        return 0
    extent = max(physical_range[1] - physical_range[0], physical_range[3] - physical_range[2],
                 physical_range[5] - physical_range[4])
    return int(max(15.0 / frequency, 10.0 * extent / C0) / dt)


def window_weight(n, start, end, window):
    if n < start or n >= end:
        return 0.0
    if window == "hanning":
        return 0.5 * (1.0 - math.cos(2.0 * math.pi * (n - start) / (end - start)))
    return 1.0


def straddle(coords, position):
    """[(index, weight)] interpolating ``position`` between its two neighbouring
    sample planes; exact hits and positions outside the samples use one plane."""
    c = np.asarray(coords, dtype=np.float64)
    if c.size == 1:
        return [(0, 1.0)]
    hi = int(np.searchsorted(c, position, side="left"))
    if hi == 0:
        return [(0, 1.0)]
    if hi == c.size:
        return [(c.size - 1, 1.0)]
    lo = hi - 1
    tol = 1e-12 * max(abs(position), abs(c[lo]), abs(c[hi]), 1.0)
    if abs(position - c[lo]) <= tol:
        return [(lo, 1.0)]
    if abs(position - c[hi]) <= tol:
        return [(hi, 1.0)]
    t = (position - c[lo]) / (c[hi] - c[lo])
    return [(lo, 1.0 - t), (hi, t)]


def _centre_average(data, axes):
    """Average neighbouring samples along the given plane axes (0 or 1) so that
    every component lands on the plane's cell centres."""
    for a in axes:
        ax = data.ndim - 2 + a
        n = data.shape[ax]
        data = 0.5 * (np.take(data, range(0, n - 1), axis=ax) + np.take(data, range(1, n), axis=ax))
    return data


class OracleObservers:
    """Samples recorded by the oracle for every monitor of a plan."""

    def __init__(self, plan):
        prep = plan.prepared
        self.plan = plan
        self.nodes = {a: np.asarray(prep.nodes(a), dtype=np.float64) for a in "xyz"}
        self.primal = {a: np.asarray(prep.primal(a), dtype=np.float64) for a in "xyz"}
        self.physical = tuple(prep.physical_domain_range)
        self.monitors = list(plan.observers.monitors) if plan.observers is not None else []
        self.frequencies = tuple(plan.observers.frequencies) if plan.observers is not None else ()
        steps, dt = plan.time_steps, plan.dt
        window, pulsed = plan.meta.get("window", "none"), plan.meta.get("pulsed", True)
        f = np.asarray(self.frequencies, dtype=np.float64)
        w = np.zeros((steps, f.size))
        for q, fq in enumerate(f):
            s0 = start_step(fq, dt, window, pulsed, self.physical)
            for n in range(steps):
                w[n, q] = window_weight(n, s0, steps, window)
        phase = 2.0 * math.pi * f[None, :] * dt * np.arange(steps, dtype=np.float64)[:, None]
        lag = math.pi * dt * f[None, :]
        self.w = {"E": (w * np.cos(phase), w * np.sin(phase)), "H": (w * np.cos(phase - lag), w * np.sin(phase - lag))}
        self.wsum = w.sum(axis=0)
        # what to record: per monitor, per component
        self.points = {}   # (comp, (i, j, k)) -> accumulator (2, F)
        self.planes = {}   # (comp, axis, index) -> accumulator (2, F, a, b)
        self.layout = []
        for m in self.monitors:
            mf = tuple(m.frequencies) if m.frequencies else self.frequencies
            entry = {"monitor": m, "fidx": [self.frequencies.index(x) for x in mf], "freqs": mf, "comps": {}}
            for comp in m.fields:
                coords = yee_coords(self.nodes, comp)
                if m.kind == "point":
                    idx = tuple(int(np.argmin(np.abs(c - p))) for c, p in zip(coords, m.position))
                    self.points.setdefault((comp, idx), None)
                    entry["comps"][comp] = ("point", idx)
                else:
                    ai = _AX[m.axis]
                    pos = float(m.position)
                    if m.compute_flux:
                        nd = self.nodes[m.axis]
                        pos = float(nd[int(np.argmin(np.abs(nd - pos)))])
                    samples = straddle(coords[ai], pos)
                    for index, _ in samples:
                        self.planes.setdefault((comp, m.axis, index), None)
                    across = tuple(c for i, c in enumerate(coords) if i != ai)
                    entry["comps"][comp] = ("plane", samples, across)
            self.layout.append(entry)
        nf = f.size
        for key in self.points:
            self.points[key] = np.zeros((2, nf))
        for key in self.planes:
            comp, axis, _ = key
            shp = [s for i, s in enumerate(plan.field_shape(comp)) if i != _AX[axis]]
            self.planes[key] = np.zeros((2, nf) + tuple(shp))

    def record(self, n, fields):
        """Accumulate the samples of the state after step ``n``."""
        for (comp, idx), acc in self.points.items():
            wc, ws = self.w[comp[0]]
            v = float(fields[comp][idx])
            acc[0] += v * wc[n]
            acc[1] += v * ws[n]
        for (comp, axis, index), acc in self.planes.items():
            wc, ws = self.w[comp[0]]
            sl = np.take(fields[comp], index, axis=_AX[axis]).astype(np.float64)
            acc[0] += wc[n][:, None, None] * sl[None]
            acc[1] += ws[n][:, None, None] * sl[None]

    def results(self):
        scale = np.where(self.wsum > 0, 2.0 / np.where(self.wsum > 0, self.wsum, 1.0), 0.0)
        out = {}
        for e in self.layout:
            m, fidx = e["monitor"], e["fidx"]
            single = len(fidx) == 1
            comps = {}
            for comp, spec in e["comps"].items():
                if spec[0] == "point":
                    acc = self.points[(comp, spec[1])]
                    val = (acc[0, fidx] + 1j * acc[1, fidx]) * scale[fidx]
                    comps[comp] = val[0] if single else val
                else:
                    _, samples, across = spec
                    val = 0.0
                    for index, weight in samples:
                        acc = self.planes[(comp, m.axis, index)]
                        val = val + weight * (acc[0, fidx] + 1j * acc[1, fidx]) * scale[fidx][:, None, None]
                    comps[comp] = {"data": val[0] if single else val, "coords": across}
            payload = {"fields": m.fields, "frequencies": e["freqs"], "components": comps}
            if m.kind != "point" and m.compute_flux:
                payload["flux"] = self._flux(m, comps, single)
            out[m.name] = payload
        return out

    def _flux(self, m, comps, single):
        """Normal Poynting flux through the physical part of a flux plane."""
        a_name, b_name = _PLANE_AXES[m.axis]
        ai, bi = _AX[a_name], _AX[b_name]
        # every tangential component averaged onto the plane's cell centres: a
        # component on nodes along a plane axis is averaged along that axis
        fields = {}
        for comp, v in comps.items():
            data = np.asarray(v["data"])
            data = data[None] if single else data
            on_nodes = [k for k, ax in enumerate((a_name, b_name))
                        if not ((comp[0] == "E") == (ax == comp[1].lower()))]
            fields[comp] = _centre_average(data, on_nodes)
        half = {a: 0.5 * (self.nodes[a][1:] + self.nodes[a][:-1]) for a in "xyz"}
        ca, cb = half[a_name], half[b_name]
        lo = {a: self.physical[2 * _AX[a]] for a in "xyz"}
        hi = {a: self.physical[2 * _AX[a] + 1] for a in "xyz"}

        def keep(c, a):
            tol = 1e-7 * max(1.0, float(np.max(np.abs(c))), abs(lo[a]), abs(hi[a]))
            return np.flatnonzero((c >= lo[a] - tol) & (c <= hi[a] + tol))
        ka, kb = keep(ca, a_name), keep(cb, b_name)
        area = self.primal[a_name][ka][:, None] * self.primal[b_name][kb][None, :]
        # S.n with (n, b, c) cyclic: E_b H_c* - E_c H_b*
        n_ax = m.axis
        b_ax, c_ax = "xyz"[(_AX[n_ax] + 1) % 3], "xyz"[(_AX[n_ax] + 2) % 3]

        def f(name):
            return fields[name][:, ka][:, :, kb]
        s = f("E" + b_ax) * np.conj(f("H" + c_ax)) - f("E" + c_ax) * np.conj(f("H" + b_ax))
        sign = 1.0 if m.normal_direction == "+" else -1.0
        flux = (0.5 * sign * np.real(s) * area).sum(axis=(-2, -1))
        del ai, bi
        return float(flux[0]) if single else flux
