"""Source compilation: point-dipole patches and soft plane-wave surface terms.

* Point dipole (``_prepare_point_dipole_source``, fdtd/excitation/injection.py:214-389):
  patch = -dt * pol * profile / eps_edge, profile either the normalised Gaussian
  (``_normalized_point_dipole_profile`` :36-41, control volumes :44-60) or the
  trilinear "ideal" stencil (``_ideal_point_dipole_term`` :127-202).  Applied each
  step as E[patch] += s(n*dt) * patch right after the E update
  (fdtd/runtime/stepping.py:2309-2310).
* Soft plane wave (``_prepare_power_normalized_surface_source``, injection.py:807-968):
  the low-face discrete TFSF specs (fdtd/excitation/tfsf_specs.py:46-303) on the
  launch plane at 5% of the physical span (spatial.py:64-76), phase speed from the
  Yee dispersion relation (fdtd/dispersion.py:6-48), unit-power scale
  (injection.py:732-765) divided by sqrt(cos(k_axis*d/2)).  H terms are applied at
  t after the H update, E terms at t + dt/2 after the E update
  (stepping.py:1776-1777, :2300-2307).
A CW waveform is evaluated through the same time-shifted form, which equals the
reference's cos/sin-phased patch pair algebraically.
"""

from __future__ import annotations

import math

import numpy as np

from ..constants import C_0, ETA_0
from ..sources import POINT_DIPOLE_REFERENCE_WIDTH, PlaneWave, PointDipole, source_time_dict

_AX = {"x": 0, "y": 1, "z": 2}


def _control_widths(coords):
    coords = np.asarray(coords, dtype=np.float64)
    n = coords.size
    if n <= 1:
        return np.ones_like(coords)
    d = np.diff(coords)
    w = np.empty(n)
    w[0] = 0.5 * d[0]
    w[-1] = 0.5 * d[-1]
    if n > 2:
        w[1:-1] = 0.5 * (d[:-1] + d[1:])
    return w


def _axis_index_window(nodes, lo_value, hi_value, cap):
    if hi_value < nodes[0] or lo_value > nodes[-1]:
        return 0, 0
    lo = max(0, int(np.searchsorted(nodes, lo_value, side="right")) - 1)
    hi = min(int(cap), int(np.searchsorted(nodes, hi_value, side="left")) + 1)
    return lo, hi


def _ideal_axis_weights(coords, position):
    n = coords.size
    if n == 1:
        return [0], [1.0]
    upper = int(np.searchsorted(coords, position, side="left"))
    if upper <= 0:
        return [0], [1.0]
    if upper >= n:
        return [n - 1], [1.0]
    lower = upper - 1
    span = coords[upper] - coords[lower]
    if span <= np.finfo(np.float32).eps:
        return [lower], [1.0]
    frac = (position - coords[lower]) / span
    if frac <= 1e-12:
        return [lower], [1.0]
    if frac >= 1.0 - 1e-12:
        return [upper], [1.0]
    return [lower, upper], [1.0 - frac, frac]


def point_dipole_terms(prepared, source: PointDipole, eps_edges, dt, st):
    """Additive patch terms of one dipole: list of dicts(field, offset, patch, st)."""
    terms = []
    cutoff = 3.0 * POINT_DIPOLE_REFERENCE_WIDTH if source.profile == "ideal" else \
        3.0 * max(source.width, 0.5 * POINT_DIPOLE_REFERENCE_WIDTH)
    sx, sy, sz = source.position
    nx, ny, nz = prepared.shape
    x, y, z = prepared.x_nodes64, prepared.y_nodes64, prepared.z_nodes64
    i0, i1 = _axis_index_window(x, sx - cutoff, sx + cutoff, nx)
    j0, j1 = _axis_index_window(y, sy - cutoff, sy + cutoff, ny)
    k0, k1 = _axis_index_window(z, sz - cutoff, sz + cutoff, nz)
    if i1 <= i0 or j1 <= j0 or k1 <= k0:
        return terms
    for comp, pol in zip(("Ex", "Ey", "Ez"), source.polarization):
        if np.isclose(pol, 0.0):
            continue
        cx, cy, cz = prepared.component_coords(comp)
        vol = _control_widths(cx)[:, None, None] * _control_widths(cy)[None, :, None] * _control_widths(cz)[None, None, :]
        eps = eps_edges[comp]
        a0, a1, b0, b1, c0, c1 = i0, i1, j0, j1, k0, k1
        if comp == "Ex":
            a1 = min(i1, nx - 1)
        elif comp == "Ey":
            b1 = min(j1, ny - 1)
        else:
            c1 = min(k1, nz - 1)
        if a1 <= a0 or b1 <= b0 or c1 <= c0:
            continue
        px, py, pz = cx[a0:a1], cy[b0:b1], cz[c0:c1]
        if source.profile == "ideal":
            xi, xw = _ideal_axis_weights(px, sx)
            yi, yw = _ideal_axis_weights(py, sy)
            zi, zw = _ideal_axis_weights(pz, sz)
            off = (a0 + xi[0], b0 + yi[0], c0 + zi[0])
            shape = (xi[-1] - xi[0] + 1, yi[-1] - yi[0] + 1, zi[-1] - zi[0] + 1)
            patch = np.zeros(shape)
            scale = -dt * pol
            for ii, wx in zip(xi, xw):
                for jj, wy in zip(yi, yw):
                    for kk, wz in zip(zi, zw):
                        g = (a0 + ii, b0 + jj, c0 + kk)
                        patch[ii - xi[0], jj - yi[0], kk - zi[0]] += scale * wx * wy * wz / (eps[g] * vol[g])
        else:
            dist2 = (px[:, None, None] - sx) ** 2 + (py[None, :, None] - sy) ** 2 + (pz[None, None, :] - sz) ** 2
            prof = np.exp(-dist2 / (2.0 * source.width ** 2))
            mass = max(float(np.sum(prof * vol[a0:a1, b0:b1, c0:c1])), float(np.finfo(np.float32).eps))
            patch = (-dt * pol / eps[a0:a1, b0:b1, c0:c1]) * (prof / mass)
            off = (a0, b0, c0)
        terms.append({"field": comp, "offset": off, "patch": patch, "st": st})
    return terms


def solve_numerical_wavenumber(omega, dt, direction, deltas):
    """Bisection solve of the 3-D Yee dispersion relation for |k| (fdtd/dispersion.py:6-48)."""
    if omega <= 0.0:
        return 0.0
    target = math.sin(0.5 * omega * dt) ** 2 / (C_0 ** 2 * dt ** 2)
    if target <= 0.0:
        return omega / C_0
    limits = []
    for axis, comp in zip("xyz", direction):
        a = abs(float(comp))
        if a <= 1e-12:
            continue
        limits.append(2.0 * math.pi / max(a * float(deltas[axis]), 1e-30))
    if not limits:
        return omega / C_0

    def residual(k):
        tot = 0.0
        for axis, comp in zip("xyz", direction):
            a = abs(float(comp))
            if a <= 1e-12:
                continue
            d = float(deltas[axis])
            tot += math.sin(0.5 * k * a * d) ** 2 / (d * d)
        return tot - target

    lo, hi = 0.0, min(limits)
    for _ in range(80):
        mid = 0.5 * (lo + hi)
        if residual(mid) >= 0.0:
            hi = mid
        else:
            lo = mid
    return 0.5 * (lo + hi)


def _physical_interior(prepared, axis):
    size = prepared.shape[_AX[axis]]
    lo = prepared.pml_layers[axis][0]
    hi = size - prepared.pml_layers[axis][1] - 1
    if hi <= lo:
        raise ValueError(f"Scene has no physical interior along {axis}.")
    return lo, hi


def soft_plane_wave_index(prepared, axis, direction_component, fraction=0.05):
    """Launch plane at 5% of the physical span from the entry side (spatial.py:64-76)."""
    nodes = prepared.nodes(axis)
    lo, hi = _physical_interior(prepared, axis)
    span = nodes[hi] - nodes[lo]
    target = nodes[lo] + fraction * span if direction_component >= 0.0 else nodes[hi] - fraction * span
    # the reference argmins over float32 node coordinates
    n32 = nodes.astype(np.float32)
    return int(np.argmin(np.abs(n32 - np.float32(target))))


def _region_spacing(prepared, injection_axis, plane_index, direction_sign):
    """Per-axis spacing local to the launch footprint (spatial.py:93-131)."""
    deltas = {}
    for axis in "xyz":
        cells = prepared.primal(axis)
        lo, hi = _physical_interior(prepared, axis)
        interior = cells[lo:hi] if hi > lo else cells
        if (interior.max() - interior.min()) <= 1e-9 * max(interior.max(), 1e-300):
            deltas[axis] = float(interior.min())
        elif axis == injection_axis:
            ci = plane_index if direction_sign > 0 else plane_index - 1
            deltas[axis] = float(cells[min(max(ci, 0), len(cells) - 1)])
        else:
            deltas[axis] = float(interior.mean())
    return deltas


def _magnetic_unit_vector(d, p):
    mx = -(d[1] * p[2] - d[2] * p[1])
    my = -(d[2] * p[0] - d[0] * p[2])
    mz = -(d[0] * p[1] - d[1] * p[0])
    mag = math.sqrt(mx * mx + my * my + mz * mz)
    if mag <= 1e-12:
        raise ValueError("Plane-wave magnetic direction is ill-defined.")
    return {"Hx": mx / mag, "Hy": my / mag, "Hz": mz / mag}


def surface_face_specs(lower, upper, axis, side):
    """(electric, magnetic) one-cell curl-correction specs of one TFSF face
    (tfsf_specs.py:46-303)."""
    ix0, iy0, iz0 = lower
    ix1, iy1, iz1 = upper
    def spec(field, incident, offsets, shape, sample_offsets, delta_axis, sign):
        return {"field": field, "incident": incident, "offsets": offsets, "shape": shape,
                "sample_offsets": sample_offsets, "delta_axis": delta_axis, "sign": float(sign)}
    if axis == "x":
        xe = ix0 if side == "low" else ix1
        xs_e = ix0 - 1 if side == "low" else ix1
        xh = ix0 - 1 if side == "low" else ix1
        s = 1.0 if side == "low" else -1.0
        electric = [spec("Ey", "Hz", (xe, iy0, iz0), (1, iy1 - iy0, iz1 - iz0 + 1), (xs_e, iy0, iz0), "x", +s),
                    spec("Ez", "Hy", (xe, iy0, iz0), (1, iy1 - iy0 + 1, iz1 - iz0), (xs_e, iy0, iz0), "x", -s)]
        magnetic = [spec("Hy", "Ez", (xh, iy0, iz0), (1, iy1 - iy0 + 1, iz1 - iz0), (xe, iy0, iz0), "x", +s),
                    spec("Hz", "Ey", (xh, iy0, iz0), (1, iy1 - iy0, iz1 - iz0 + 1), (xe, iy0, iz0), "x", -s)]
    elif axis == "y":
        ye = iy0 if side == "low" else iy1
        ys_e = iy0 - 1 if side == "low" else iy1
        yh = iy0 - 1 if side == "low" else iy1
        s = 1.0 if side == "low" else -1.0
        electric = [spec("Ex", "Hz", (ix0, ye, iz0), (ix1 - ix0, 1, iz1 - iz0 + 1), (ix0, ys_e, iz0), "y", -s),
                    spec("Ez", "Hx", (ix0, ye, iz0), (ix1 - ix0 + 1, 1, iz1 - iz0), (ix0, ys_e, iz0), "y", +s)]
        magnetic = [spec("Hx", "Ez", (ix0, yh, iz0), (ix1 - ix0 + 1, 1, iz1 - iz0), (ix0, ye, iz0), "y", -s),
                    spec("Hz", "Ex", (ix0, yh, iz0), (ix1 - ix0, 1, iz1 - iz0 + 1), (ix0, ye, iz0), "y", +s)]
    else:
        ze = iz0 if side == "low" else iz1
        zs_e = iz0 - 1 if side == "low" else iz1
        zh = iz0 - 1 if side == "low" else iz1
        s = 1.0 if side == "low" else -1.0
        electric = [spec("Ex", "Hy", (ix0, iy0, ze), (ix1 - ix0, iy1 - iy0 + 1, 1), (ix0, iy0, zs_e), "z", +s),
                    spec("Ey", "Hx", (ix0, iy0, ze), (ix1 - ix0 + 1, iy1 - iy0, 1), (ix0, iy0, zs_e), "z", -s)]
        magnetic = [spec("Hx", "Ey", (ix0, iy0, zh), (ix1 - ix0 + 1, iy1 - iy0, 1), (ix0, iy0, ze), "z", +s),
                    spec("Hy", "Ex", (ix0, iy0, zh), (ix1 - ix0, iy1 - iy0 + 1, 1), (ix0, iy0, ze), "z", -s)]
    return electric, magnetic


def component_positions(prepared, comp, offsets, shape):
    cx, cy, cz = prepared.component_coords(comp)
    px = cx[offsets[0]:offsets[0] + shape[0]]
    py = cy[offsets[1]:offsets[1] + shape[1]]
    pz = cz[offsets[2]:offsets[2] + shape[2]]
    return np.stack(np.broadcast_arrays(px[:, None, None], py[None, :, None], pz[None, None, :]), axis=-1)


def _coeff_patch(prepared, spec, coef):
    """curl coefficient slice divided by the spacing along the derivative axis
    (tfsf_common.py:316-334)."""
    off, shp, axis = spec["offsets"], spec["shape"], spec["delta_axis"]
    spacing = prepared.dual(axis) if spec["field"].startswith("E") else prepared.primal(axis)
    delta = float(spacing[int(off[_AX[axis]])])
    curl = coef[spec["field"]][1]
    return curl[off[0]:off[0] + shp[0], off[1]:off[1] + shp[1], off[2]:off[2] + shp[2]] / delta


def soft_plane_wave_terms(prepared, source: PlaneWave, coef, dt, st):
    """(electric_terms, magnetic_terms) of a soft plane wave."""
    d = source.direction
    axis = source.injection_axis or "xyz"[int(np.argmax(np.abs(d)))]
    ai = _AX[axis]
    sign = 1 if d[ai] >= 0.0 else -1
    plane = soft_plane_wave_index(prepared, axis, d[ai])
    nx, ny, nz = prepared.shape
    lower, upper, aperture = [], [], []
    comp_range = prepared.domain_range
    for idx, ax in enumerate("xyz"):
        if ax == axis:
            lo, hi = (plane, plane + 1) if sign > 0 else (plane - 1, plane)
            pc = float(prepared.nodes(ax)[plane])
            aperture.append((pc, pc))
        else:
            lo, hi = 0, (nx, ny, nz)[idx] - 1
            aperture.append((comp_range[2 * idx], comp_range[2 * idx + 1]))
        lower.append(lo)
        upper.append(hi)
    electric_specs, magnetic_specs = surface_face_specs(tuple(lower), tuple(upper), axis,
                                                        "low" if sign > 0 else "high")
    omega = 2.0 * math.pi * st["frequency"]
    deltas = _region_spacing(prepared, axis, plane, sign)
    k_num = solve_numerical_wavenumber(omega, dt, d, deltas)
    speed = omega / k_num if k_num > 1e-12 else C_0
    extents = [aperture[i][1] - aperture[i][0] for i in range(3) if i != ai]
    unit_power = extents[0] * extents[1] * abs(d[ai]) / (2.0 * ETA_0)
    power = 1.0 / math.sqrt(unit_power)
    factor = math.cos(0.5 * k_num * d[ai] * deltas[axis])
    if factor <= 0.0:
        raise ValueError("PlaneWave grid is too coarse for positive discrete normal power.")
    power /= math.sqrt(factor)
    e_vec = {"Ex": power * source.polarization[0], "Ey": power * source.polarization[1],
             "Ez": power * source.polarization[2]}
    h_vec = {k: power * v / ETA_0 for k, v in _magnetic_unit_vector(d, source.polarization).items()}
    center = list(prepared.domain.center)
    center[ai] = float(prepared.nodes(axis)[plane])
    ref = np.asarray(center)
    dvec = np.asarray(d)

    def build(specs, vec):
        out = []
        for sp in specs:
            cs = vec[sp["incident"]]
            if abs(sp["sign"] * cs) <= 1e-12:
                continue
            cp = _coeff_patch(prepared, sp, coef)
            pos = component_positions(prepared, sp["incident"], sp["sample_offsets"], sp["shape"])
            delay = np.sum((pos - ref) * dvec, axis=-1) / speed
            scale = sp["sign"] * cs * cp
            if np.max(np.abs(scale)) <= 1e-12:
                continue
            out.append({"field": sp["field"], "offset": sp["offsets"], "patch": scale, "delay": delay, "st": st})
        return out

    return build(electric_specs, h_vec), build(magnetic_specs, e_vec)


def primary_source_time(source, default_frequency):
    st = source.source_time
    if st is None:
        from ..sources import CW
        st = CW(frequency=default_frequency)
    return source_time_dict(st)
