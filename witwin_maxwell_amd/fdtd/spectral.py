"""Running-DFT schedules, observer layout and monitor result assembly.

* Windows / start steps: ``resolve_spectral_window_type``, ``compute_spectral_start_step``
  and ``compute_window_weight`` (fdtd/runtime/spectral.py:227-297): pulsed sources
  force window "none" starting at step 0.
* Phase: step n carries (cos, sin)(omega*n*dt) with a + sign, applied to the fields
  after the whole step (spectral.py:558-611, stepping.py:2313-2330).  H samples are
  retarded by -omega*dt/2 (``observer_stagger_weights``, spectral.py:196-220).
  The weights here are evaluated directly in float64 for every step, instead of
  the reference's float32 phase recurrence, which drifts by ~T*eps_fp32.
* Point observers take the nearest component sample (observers.py:275-282); plane
  observers interpolate between the two straddling planes (:296-388); flux planes
  snap to the nearest node plane (:390-405).
* Results: scale 2/sum(w) (observers.py:236-241), plane alignment onto cell centres
  (:25-37, :126-200) and flux = sum 1/2 Re(E x H*).n dA with exact cell widths
  (fdtd/quadrature.py:78-110, :185-233).
"""

from __future__ import annotations

import numpy as np

from ..monitors import FluxMonitor, PlaneMonitor, PointMonitor

_PLANE_ALIGNMENT_RULES = {
    "x": {"coord_sources": {"y": ("Ey", 0, "Hz", 0), "z": ("Ez", 1, "Hy", 1)},
          "average_axis": {"Ex": (0, 1), "Ey": 1, "Ez": 0, "Hy": 0, "Hz": 1}},
    "y": {"coord_sources": {"x": ("Ex", 0, "Hz", 0), "z": ("Ez", 1, "Hx", 1)},
          "average_axis": {"Ex": 1, "Ey": (0, 1), "Ez": 0, "Hx": 0, "Hz": 1}},
    "z": {"coord_sources": {"x": ("Ex", 0, "Hy", 0), "y": ("Ey", 1, "Hx", 1)},
          "average_axis": {"Ex": 1, "Ey": 0, "Ez": (0, 1), "Hx": 0, "Hy": 1}},
}
_COORD_NAMES = {"x": ("y", "z"), "y": ("x", "z"), "z": ("x", "y")}


def spectral_start_step(prepared, frequency, dt, window, pulsed):
    if pulsed or window == "none":
        return 0
    from ..constants import C_0
    period = 1.0 / frequency
    r = prepared.physical_domain_range
    size = max(r[1] - r[0], r[3] - r[2], r[5] - r[4])
    transient = max(15 * period, 2 * size / C_0 * 5)
    return int(transient / dt)


def window_weight(n, start, end, window):
    if n < start or n >= end:
        return 0.0
    pos = (n - start) / (end - start)
    if window == "hanning":
        return 0.5 * (1.0 - np.cos(2 * np.pi * pos))
    if window == "ramp":
        if pos < 0.1:
            return 0.5 * (1.0 - np.cos(np.pi * pos / 0.1))
        return 1.0
    return 1.0


def weight_tables(frequencies, dt, steps, starts, window):
    """(E_cos, E_sin, H_cos, H_sin) of shape (steps, F) and the per-frequency
    window sums; every entry already includes the window weight."""
    freqs = np.asarray(frequencies, dtype=np.float64)
    nf = freqs.size
    n = np.arange(steps, dtype=np.float64)[:, None]
    w = np.zeros((steps, nf))
    for f in range(nf):
        for s in range(steps):
            w[s, f] = window_weight(s, starts[f], steps, window)
    omega = 2.0 * np.pi * freqs[None, :]
    phase = omega * n * dt
    e_cos = w * np.cos(phase)
    e_sin = w * np.sin(phase)
    h_phase = phase - 0.5 * omega * dt
    h_cos = w * np.cos(h_phase)
    h_sin = w * np.sin(h_phase)
    return e_cos, e_sin, h_cos, h_sin, w.sum(axis=0), (w != 0).sum(axis=0)


def _plane_samples(coords, position):
    coords = np.asarray(coords, dtype=np.float64)
    if coords.size == 1:
        return [(0, float(coords[0]), 1.0)]
    upper = int(np.searchsorted(coords, position, side="left"))
    if upper <= 0:
        return [(0, float(coords[0]), 1.0)]
    if upper >= coords.size:
        return [(coords.size - 1, float(coords[-1]), 1.0)]
    lower = upper - 1
    lc, uc = float(coords[lower]), float(coords[upper])
    tol = 1e-12 * max(abs(position), abs(lc), abs(uc), 1.0)
    if abs(position - lc) <= tol:
        return [(lower, lc, 1.0)]
    if abs(position - uc) <= tol:
        return [(upper, uc, 1.0)]
    span = uc - lc
    if span <= tol:
        nearest = lower if abs(position - lc) <= abs(position - uc) else upper
        return [(nearest, float(coords[nearest]), 1.0)]
    wu = (position - lc) / span
    return [(lower, lc, 1.0 - wu), (upper, uc, wu)]


class ObserverLayout:
    """Groups monitors into point samples and (component, axis, index) planes."""

    def __init__(self, prepared, monitors, default_frequencies):
        self.prepared = prepared
        self.monitors = [m for m in monitors if isinstance(m, (PointMonitor, PlaneMonitor))]
        freqs = []
        for f in list(default_frequencies or ()):
            if f not in freqs:
                freqs.append(float(f))
        for m in self.monitors:
            for f in (m.frequencies or ()):
                if f not in freqs:
                    freqs.append(float(f))
        self.frequencies = tuple(freqs)
        self.points = []       # (component, (i,j,k))
        self.planes = []       # (component, axis, plane_index)
        self.records = []      # per monitor
        for m in self.monitors:
            mf = tuple(m.frequencies) if m.frequencies else tuple(default_frequencies or ())
            if not mf:
                raise ValueError(f"monitor {m.name!r} has no frequencies")
            fidx = tuple(self.frequencies.index(f) for f in mf)
            rec = {"monitor": m, "frequencies": mf, "freq_index": fidx, "components": {}}
            for comp in m.fields:
                coords = prepared.component_coords(comp)
                if isinstance(m, PointMonitor):
                    idx = tuple(int(np.argmin(np.abs(c - p))) for c, p in zip(coords, m.position))
                    key = (comp, idx)
                    if key not in self.points:
                        self.points.append(key)
                    rec["components"][comp] = {"point": self.points.index(key), "field_index": idx}
                else:
                    axis = m.axis
                    ai = "xyz".index(axis)
                    position = m.position
                    if m.compute_flux:
                        nodes = prepared.nodes(axis)
                        position = float(nodes[int(np.argmin(np.abs(nodes - position)))])
                    samples = []
                    for plane_index, plane_pos, weight in _plane_samples(coords[ai], position):
                        key = (comp, axis, plane_index)
                        if key not in self.planes:
                            self.planes.append(key)
                        samples.append((self.planes.index(key), weight, plane_pos))
                    other = [c for i, c in enumerate(coords) if i != ai]
                    rec["components"][comp] = {"samples": samples, "coords": tuple(other),
                                               "plane_index": max(samples, key=lambda s: s[1])[0]}
            self.records.append(rec)

    def plane_shape(self, shapes, key):
        comp, axis, _ = key
        s = shapes[comp]
        return {"x": (s[1], s[2]), "y": (s[0], s[2]), "z": (s[0], s[1])}[axis]


def _avg_adjacent(data, axis):
    axis = axis + (data.ndim - 2)
    lo = [slice(None)] * data.ndim
    hi = [slice(None)] * data.ndim
    lo[axis] = slice(0, -1)
    hi[axis] = slice(1, None)
    return 0.5 * (data[tuple(lo)] + data[tuple(hi)])


def _exact_widths(prepared, axis, points):
    centers = prepared.half(axis)
    widths = prepared.primal(axis)
    idx = np.array([int(np.argmin(np.abs(centers - v))) for v in np.asarray(points).reshape(-1)], dtype=np.int64)
    return widths[idx]


def _is_torch(x):
    return type(x).__module__.startswith("torch")


def _complex(re, im):
    if _is_torch(re):
        import torch
        return torch.complex(re.double(), im.double())
    return re + 1j * im


def _scale_like(scale, ref):
    if _is_torch(ref):
        import torch
        return torch.as_tensor(scale, dtype=torch.float64, device=ref.device)
    return scale


def assemble_monitor_results(layout: ObserverLayout, point_re, point_im, plane_re, plane_im, window_sum):
    """Monitor payloads from accumulated sums.

    ``point_re/im``: (F, n_points) arrays; ``plane_re/im``: list of (F, a, b)
    arrays per plane; ``window_sum``: (F,) window normalisation.  The sums may be
    numpy arrays or torch tensors; with tensors every payload (and the flux) is a
    complex128 / float64 tensor that keeps the autograd graph of the sums, which
    is how monitor losses reach the adjoint."""
    prepared = layout.prepared
    scale = np.where(window_sum > 0, 2.0 / np.where(window_sum > 0, window_sum, 1.0), 0.0)
    results = {}
    for rec in layout.records:
        m = rec["monitor"]
        fidx = list(rec["freq_index"])
        single = len(fidx) == 1
        comps = {}
        for comp, info in rec["components"].items():
            if "point" in info:
                p = info["point"]
                data = _complex(point_re[fidx, p], point_im[fidx, p]) * _scale_like(scale[fidx], point_re)
                comps[comp] = data[0] if single else data
            else:
                acc = None
                for plane, weight, _ in info["samples"]:
                    sc = _scale_like(scale[fidx][:, None, None], plane_re[plane])
                    val = _complex(plane_re[plane][fidx], plane_im[plane][fidx]) * sc
                    acc = weight * val if acc is None else acc + weight * val
                comps[comp] = {"data": acc[0] if single else acc, "coords": info["coords"]}
        entry = {"kind": "point" if isinstance(m, PointMonitor) else "plane", "fields": m.fields,
                 "frequency": rec["frequencies"][0], "frequencies": rec["frequencies"], "components": comps}
        if isinstance(m, PointMonitor):
            entry["position"] = m.position
            for comp, v in comps.items():
                entry[comp] = v
            if len(comps) == 1:
                (comp, v), = comps.items()
                entry["component"] = comp
                entry["data"] = v
                entry["field_index"] = rec["components"][comp]["field_index"]
        else:
            entry.update({"axis": m.axis, "position": m.position, "compute_flux": m.compute_flux,
                          "normal_direction": m.normal_direction})
            if len(comps) == 1:
                (comp, v), = comps.items()
                entry["component"] = comp
                entry["data"] = v["data"]
                entry["coords"] = v["coords"]
                entry[comp] = v["data"]
            else:
                aligned = _align_plane(m.axis, comps)
                if aligned is not None:
                    if m.compute_flux:
                        aligned = _crop_to_physical(prepared, m.axis, aligned)
                    a_name, b_name = _COORD_NAMES[m.axis]
                    entry[a_name] = aligned[a_name]
                    entry[b_name] = aligned[b_name]
                    entry["coords"] = (aligned[a_name], aligned[b_name])
                    entry["cell_widths"] = {a_name: _exact_widths(prepared, a_name, aligned[a_name]),
                                            b_name: _exact_widths(prepared, b_name, aligned[b_name])}
                    for comp, v in aligned["fields"].items():
                        entry[comp] = v
                    if m.compute_flux:
                        entry["flux"] = plane_flux(entry, single)
                        entry["power"] = entry["flux"]
        results[m.name] = entry
    return results


def _align_plane(axis, comps):
    rules = _PLANE_ALIGNMENT_RULES[axis]
    a_name, b_name = _COORD_NAMES[axis]
    fields = {}
    for comp, payload in comps.items():
        avg = rules["average_axis"].get(comp)
        if avg is None:
            continue
        data = payload["data"]
        for ax in ((avg,) if isinstance(avg, int) else avg):
            data = _avg_adjacent(data, ax)
        fields[comp] = data

    def coord(name):
        src = rules["coord_sources"][name]
        for comp, ci in zip(src[::2], src[1::2]):
            if comp in comps:
                return np.asarray(comps[comp]["coords"][ci])
        return None
    ca, cb = coord(a_name), coord(b_name)
    if ca is None or cb is None:
        return None
    return {a_name: ca, b_name: cb, "fields": fields}


def _crop_to_physical(prepared, axis, aligned):
    a_name, b_name = _COORD_NAMES[axis]
    r = prepared.physical_domain_range
    bounds = {"x": (r[0], r[1]), "y": (r[2], r[3]), "z": (r[4], r[5])}
    masks = []
    for name in (a_name, b_name):
        vals = np.asarray(aligned[name], dtype=np.float64)
        lo, hi = bounds[name]
        tol = 1e-7 * max(1.0, float(np.max(np.abs(vals))), abs(lo), abs(hi))
        masks.append(np.flatnonzero((vals >= lo - tol) & (vals <= hi + tol)))
    ia, ib = masks
    if ia.size == 0 or ib.size == 0:
        raise ValueError("FluxMonitor plane has no samples inside the physical domain.")
    def take(v):
        if _is_torch(v):
            import torch
            return v.index_select(-2, torch.as_tensor(ia, device=v.device)).index_select(
                -1, torch.as_tensor(ib, device=v.device))
        return np.take(np.take(v, ia, axis=-2), ib, axis=-1)
    fields = {c: take(v) for c, v in aligned["fields"].items()}
    return {a_name: np.asarray(aligned[a_name])[ia], b_name: np.asarray(aligned[b_name])[ib], "fields": fields}


def plane_flux(entry, single):
    axis = entry["axis"]
    ai = "xyz".index(axis)
    a_name, b_name = _COORD_NAMES[axis]
    weights = entry["cell_widths"][a_name][:, None] * entry["cell_widths"][b_name][None, :]
    direction = 1.0 if entry.get("normal_direction", "+") == "+" else -1.0
    # normal component of E x conj(H): (E_b H_c* - E_c H_b*) with (ai, b, c) cyclic
    b, c = "xyz"[(ai + 1) % 3], "xyz"[(ai + 2) % 3]
    def comp(name):
        return entry.get(name)
    def prod(e, h):
        if e is None or h is None:
            return 0.0
        return e * (h.conj() if _is_torch(h) else np.conj(h))
    cross = prod(comp("E" + b), comp("H" + c)) - prod(comp("E" + c), comp("H" + b))
    if _is_torch(cross):
        import torch
        w = torch.as_tensor(weights, dtype=torch.float64, device=cross.device)
        flux = (0.5 * direction * cross.real * w).sum(dim=(-2, -1))
        return flux[0] if single and flux.ndim else flux
    s = 0.5 * np.real(cross) * direction
    flux = np.sum(s * weights, axis=(-2, -1))
    return float(flux) if single else flux


def frequency_solution(frequencies, e_re, e_im, window_sum):
    """Full-field DFT output (Re + i Im) * 2/sum(w) (fdtd/postprocess.py:8-32)."""
    scale = np.where(window_sum > 0, 2.0 / np.where(window_sum > 0, window_sum, 1.0), 0.0)
    out = {}
    for comp in e_re:
        out[comp] = (e_re[comp] + 1j * e_im[comp]) * scale.reshape((-1,) + (1,) * (e_re[comp].ndim - 1))
    return out
