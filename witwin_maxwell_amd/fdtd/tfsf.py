"""Total-field/scattered-field plane-wave injection from a 1-D auxiliary line.

Restates the reference's axis-aligned auxiliary-line provider
(fdtd/excitation/tfsf_state.py:265-332 box, :668-707 non-periodic slab), the one a
normally incident ``PlaneWave`` with ``TFSF(bounds)`` or ``TFSF.slab(axis=...)``
selects:

* TF/SF region indices: nearest float32 node to each bound
  (tfsf_common.py:13-23); a slab spans the whole non-PML transverse extent
  (tfsf_state.py:130-168).
* Faces: the discrete one-cell curl-correction specs of every face (box) or of
  the two faces normal to the slab axis (tfsf_specs.py:46-332).  A correction
  adds sign * vector[incident] * (curl coefficient / spacing) * line sample,
  vector = polarisation for H faces and -(d x p)/|d x p| for E faces
  (tfsf_specs.py:335-348, tfsf_common.py:316-351); zero-vector faces are
  dropped.
* Line samples are integer indices: the grid index along the propagation
  axis, mirrored (N-1-k for E, N-2-k for H) for a negative direction
  (tfsf_specs.py:478-490).
* The auxiliary line (excitation/spatial.py:311-469): E nodes at
  s_min + m ds over the computational domain's projection plus 20 absorber cells,
  sigma ramps 0.35 ((m+1)/20)^2 (E) and 0.35 ((m+1/2)/20)^2 (H) on the last 20
  samples, vacuum eps/mu from eta0 and c, a hard source E[0] = s(n dt) and E = 0
  at the far end (kernels/sources.cu:803-839).
* Step placement (fdtd/runtime/stepping.py:1773-1775, :2297-2299): H faces read
  line E^n after the H update, then the line's H advances; E faces read line
  H^{n+1/2} after the E update, then the line's E advances.
"""

from __future__ import annotations

import math

import numpy as np

from ..constants import C_0, EPSILON_0, MU_0
from ..sources import TFSF, PlaneWave
from .excitation import _AX, _magnetic_unit_vector, surface_face_specs

ABSORBER_CELLS = 20
_UNIFORM_RTOL = 1e-6


def _nearest_index(nodes, value):
    n32 = np.asarray(nodes, dtype=np.float64).astype(np.float32)
    return int(np.argmin(np.abs(n32 - np.float32(value))))


def _axis_direction(direction):
    axis, sign = None, 0
    for a, comp in zip("xyz", direction):
        if abs(float(comp)) > 1e-9:
            if axis is not None:
                return None
            axis, sign = a, (1 if comp > 0 else -1)
    return None if axis is None else (axis, sign)


def region_indices(prepared, injection: TFSF):
    """(lower, upper) node indices of the TF region (tfsf_state.py:130-168)."""
    shape = prepared.shape
    if injection.mode == "slab":
        axis = injection.axis
        lower, upper = [], []
        for a, size in zip("xyz", shape):
            nodes = prepared.nodes(a)
            if a == axis:
                lower.append(_nearest_index(nodes, injection.axis_bounds[0]))
                upper.append(_nearest_index(nodes, injection.axis_bounds[1]))
            else:
                lower.append(int(prepared.pml_layers[a][0]))
                upper.append(int(size - prepared.pml_layers[a][1] - 1))
        ai = _AX[axis]
        if upper[ai] <= lower[ai] + 1:
            raise ValueError(f"TFSF slab bounds must span at least two cells along {axis}.")
        if lower[ai] < prepared.pml_layers[axis][0] or upper[ai] > shape[ai] - prepared.pml_layers[axis][1] - 1:
            raise ValueError("TFSF slab bounds must lie strictly inside the non-PML simulation region.")
        return tuple(lower), tuple(upper)
    lower, upper = [], []
    for a, (lo, hi) in zip("xyz", injection.bounds):
        nodes = prepared.nodes(a)
        lower.append(_nearest_index(nodes, lo))
        upper.append(_nearest_index(nodes, hi))
    for a, lo, hi, size in zip("xyz", lower, upper, shape):
        low_margin = prepared.pml_layers[a][0] + 2
        high_margin = prepared.pml_layers[a][1] + 2
        if hi <= lo + 1:
            raise ValueError(f"TFSF bounds must span at least two cells along {a}.")
        if lo < low_margin or hi > size - high_margin - 1:
            raise ValueError("TFSF bounds must lie strictly inside the non-PML simulation region.")
    return tuple(lower), tuple(upper)


def _local_spacing(prepared, axis, start, stop):
    cells = prepared.primal(axis)
    window = cells[max(int(start), 0):min(int(stop), len(cells))]
    d_min, d_max = float(window.min()), float(window.max())
    if (d_max - d_min) > _UNIFORM_RTOL * d_max:
        raise NotImplementedError("TFSF over a graded grid region is outside this FDTD runtime")
    return d_min


def _vacuum_faces(node_eps, model, lower, upper, faces, tol=1e-5):
    for axis, side in faces:
        sl = [slice(lower[i], upper[i] + 1) for i in range(3)]
        idx = lower[_AX[axis]] if side == "low" else upper[_AX[axis]]
        sl[_AX[axis]] = slice(idx, idx + 1)
        sl = tuple(sl)
        if not np.allclose(node_eps[sl], 1.0, atol=tol, rtol=0.0) or not np.allclose(model.mu_r[sl], 1.0,
                                                                                        atol=tol, rtol=0.0):
            raise ValueError("TFSF boundary must remain in vacuum on all faces.")


def auxiliary_line(prepared, direction, ds, dt, source_values):
    """The 1-D line in float64 (excitation/spatial.py:311-401)."""
    r = prepared.domain_range
    corners = [(x, y, z) for x in r[0:2] for y in r[2:4] for z in r[4:6]]
    proj = [x * direction[0] + y * direction[1] + z * direction[2] for x, y, z in corners]
    s_min, s_max = float(min(proj)), float(max(proj)) + ABSORBER_CELLS * ds
    n_e = max(int(math.ceil((s_max - s_min) / ds)) + 1, 4)
    n_h = n_e - 1
    eta0 = math.sqrt(MU_0 / EPSILON_0)
    eps = 1.0 / (eta0 * C_0)
    mu = eta0 / C_0
    sigma_e = np.zeros(n_e)
    sigma_h = np.zeros(n_h)
    ce = min(ABSORBER_CELLS, max(n_e - 1, 0))
    ch = min(ABSORBER_CELLS, max(n_h, 0))
    if ce > 0:
        sigma_e[-ce:] = 0.35 * ((np.arange(ce) + 1.0) / ce) ** 2
    if ch > 0:
        sigma_h[-ch:] = 0.35 * ((np.arange(ch) + 0.5) / ch) ** 2
    fe = 0.5 * dt * sigma_e / eps
    fh = 0.5 * dt * sigma_h / mu
    return {"n_e": n_e, "s_min": s_min, "ds": ds,
            "decay_e": (1.0 - fe) / (1.0 + fe), "curl_e": (dt / (eps * ds)) / (1.0 + fe),
            "decay_h": (1.0 - fh) / (1.0 + fh), "curl_h": (dt / (mu * ds)) / (1.0 + fh),
            "source": np.asarray(source_values, dtype=np.float64)}


def tfsf_setup(prepared, source: PlaneWave, coef, dt, st, source_values, node_eps, model):
    """(electric line terms, magnetic line terms, auxiliary line) of a TFSF plane wave."""
    inj = source.injection
    ad = _axis_direction(source.direction)
    if ad is None:
        raise NotImplementedError("TFSF injection of an oblique plane wave is outside this FDTD runtime")
    axis, sign = ad
    slab = inj.mode == "slab"
    if slab and inj.axis != axis:
        raise ValueError(f"A TFSF slab with non-periodic transverse boundaries must be a plane wave normally "
                         f"incident along the slab axis {inj.axis!r}")
    if not slab and prepared.boundary.uses_kind("periodic"):
        raise NotImplementedError("TFSF injection currently supports only none, pml, pec, or pmc boundaries.")
    lower, upper = region_indices(prepared, inj)
    deltas = {a: _local_spacing(prepared, a, lo - 1, hi + 1) for a, lo, hi in zip("xyz", lower, upper)}
    faces = [(axis, "low"), (axis, "high")] if slab else [(a, s) for a in "xyz" for s in ("low", "high")]
    _vacuum_faces(node_eps, model, lower, upper, faces)

    d = tuple(float(c) for c in source.direction)
    p = tuple(float(c) for c in source.polarization)
    e_vec = {"Ex": p[0], "Ey": p[1], "Ez": p[2]}
    h_vec = _magnetic_unit_vector(d, p)
    ai = _AX[axis]
    n_axis = prepared.shape[ai]
    e_terms, h_terms = [], []
    for fa, side in faces:
        electric, magnetic = surface_face_specs(lower, upper, fa, side)
        for family, specs, vec, kind in (("E", electric, h_vec, "magnetic"), ("H", magnetic, e_vec, "electric")):
            for spec in specs:
                scale = vec[spec["incident"]]
                if abs(float(spec["sign"]) * scale) <= 1e-12:
                    continue
                off, shp, dax = spec["offsets"], spec["shape"], spec["delta_axis"]
                spacing = prepared.dual(dax) if family == "E" else prepared.primal(dax)
                delta = float(spacing[int(off[_AX[dax]])])
                curl = np.broadcast_to(coef[spec["field"]][1], _component_shape(spec["field"], prepared.shape))
                patch = spec["sign"] * scale * curl[off[0]:off[0] + shp[0], off[1]:off[1] + shp[1],
                                                    off[2]:off[2] + shp[2]] / delta
                start = int(spec["sample_offsets"][ai])
                length = int(shp[ai])
                if sign > 0:
                    base, step = start, 1
                else:
                    base, step = (n_axis - 1 if kind == "electric" else n_axis - 2) - start, -1
                term = {"field": spec["field"], "offset": tuple(int(v) for v in off), "patch": np.array(patch),
                        "line": kind, "line_axis": ai if length > 1 else -1, "line_base": base,
                        "line_step": step if length > 1 else 0}
                (e_terms if family == "E" else h_terms).append(term)
    aux = auxiliary_line(prepared, d, deltas[axis], dt, source_values)
    aux["lower"], aux["upper"] = lower, upper
    return e_terms, h_terms, aux


def _component_shape(comp, shape):
    nx, ny, nz = shape
    return {"Ex": (nx - 1, ny, nz), "Ey": (nx, ny - 1, nz), "Ez": (nx, ny, nz - 1),
            "Hx": (nx, ny - 1, nz - 1), "Hy": (nx - 1, ny, nz - 1), "Hz": (nx - 1, ny - 1, nz)}[comp]


def line_sample_index(term, p_local):
    """Line index read by the patch sample at local coordinates ``p_local``."""
    if term["line_axis"] < 0:
        return term["line_base"]
    return term["line_base"] + term["line_step"] * p_local[term["line_axis"]]


def advance_line_h(aux, e, h):
    """line H <- decay H - curl (E[m+1] - E[m]) (kernels/sources.cu:803-814)."""
    h[:] = aux["decay_h"] * h - aux["curl_h"] * (e[1:] - e[:-1])


def advance_line_e(aux, e, h, n):
    """line E <- decay E - curl (H[m] - H[m-1]) inside, E[0] = s(n dt), E[-1] = 0
    (kernels/sources.cu:816-839)."""
    e[1:-1] = aux["decay_e"][1:-1] * e[1:-1] - aux["curl_e"][1:-1] * (h[1:] - h[:-1])
    e[0] = aux["source"][n]
    e[-1] = 0.0
