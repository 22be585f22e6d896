"""CPML profiles and the slab (compressed) psi layout.

``cpml_profile_1d`` restates ``_cpml_profile_1d`` (fdtd/boundary/cpml.py:114-210):
physical-depth grading d in [0,1] measured from the layer interface ``thickness``
cells inward, sigma = sigma_max d^m, kappa = 1 + (kappa_max-1) d^m,
alpha = alpha_max (1-d), b = exp(-(sigma/kappa + alpha) dt / vacuum),
c = sigma (b-1) / ((sigma + alpha kappa) kappa).  ``build_cpml_profiles`` follows
:380-426 (E at nodes; H at half nodes with the high-side depth taken at
``nodes[1:]`` and sigma scaled by mu0/eps0).  ``psi_layout`` follows
``_build_cpml_memory_layout`` (:218-257).

psi is stored only inside the PML slabs (the reference's "slab" memory mode).
Outside the layers b = 1 and c = 0, so a dense psi would stay identically zero
there: the slab storage is exact, not an approximation.
"""

from __future__ import annotations

import math

import numpy as np

from ..constants import EPSILON_0, MU_0

DEFAULT_CPML_CONFIG = {"grading_order": 3.0, "kappa_max": 1.0, "alpha_max": 0.0, "reflection": 1e-6}

# psi array order used by the native layer; (field, derivative axis)
PSI_NAMES = ("psi_ex_y", "psi_ex_z", "psi_ey_x", "psi_ey_z", "psi_ez_x", "psi_ez_y",
             "psi_hx_y", "psi_hx_z", "psi_hy_x", "psi_hy_z", "psi_hz_x", "psi_hz_y")
PSI_SPECS = {"psi_ex_y": ("Ex", 1), "psi_ex_z": ("Ex", 2), "psi_ey_x": ("Ey", 0), "psi_ey_z": ("Ey", 2),
             "psi_ez_x": ("Ez", 0), "psi_ez_y": ("Ez", 1), "psi_hx_y": ("Hx", 1), "psi_hx_z": ("Hx", 2),
             "psi_hy_x": ("Hy", 0), "psi_hy_z": ("Hy", 2), "psi_hz_x": ("Hz", 0), "psi_hz_y": ("Hz", 1)}


def legacy_strength_scale(strength) -> float:
    """cpml.py:105-111: strengths <= 0 or > 100 mean 1."""
    v = float(strength)
    if v <= 0 or v > 100.0:
        return 1.0
    return v


def cpml_profile_1d(positions, nodes, thickness, dt, vacuum, sigma_scale, impedance, config, *,
                    high_positions=None, apply_low=True, apply_high=True):
    positions = np.asarray(positions, dtype=np.float64)
    nodes = np.asarray(nodes, dtype=np.float64)
    n = positions.size
    thickness = int(thickness)
    if thickness <= 0:
        return np.ones(n), np.zeros(n), np.zeros(n)

    def interface(low):
        cells = nodes.size - 1
        if thickness <= cells:
            return float(nodes[thickness]) if low else float(nodes[-1 - thickness])
        extra = thickness - cells
        if low:
            return float(nodes[-1]) + extra * float(nodes[-1] - nodes[-2])
        return float(nodes[0]) - extra * float(nodes[1] - nodes[0])

    m = float(config["grading_order"])
    kappa_max = float(config["kappa_max"])
    alpha_max = float(config["alpha_max"])
    refl = float(config["reflection"])
    numerator = -(m + 1.0) * math.log(refl) / (2.0 * float(impedance))
    hpos = positions if high_positions is None else np.asarray(high_positions, dtype=np.float64)
    depth_low = np.zeros(n)
    depth_high = np.zeros(n)
    smax_low = smax_high = 0.0
    if apply_low:
        li = interface(True)
        layer = li - float(nodes[0])
        depth_low = np.clip((li - positions) / layer, 0.0, 1.0)
        smax_low = numerator / layer * sigma_scale
    if apply_high:
        hi = interface(False)
        layer = float(nodes[-1]) - hi
        depth_high = np.clip((hpos - hi) / layer, 0.0, 1.0)
        smax_high = numerator / layer * sigma_scale
    use_high = depth_high > depth_low
    depth = np.where(use_high, depth_high, depth_low)
    peak = np.where(use_high, smax_high, smax_low)
    active = depth > 0.0
    graded = np.zeros(n)
    graded[active] = depth[active] ** m
    sigma = peak * graded
    kappa = np.ones(n)
    kappa[active] = 1.0 + (kappa_max - 1.0) * graded[active]
    alpha = np.zeros(n)
    alpha[active] = alpha_max * (1.0 - depth[active])
    b = np.exp(-(sigma / kappa + alpha) * float(dt) / float(vacuum))
    denom = sigma + alpha * kappa
    c = np.zeros(n)
    mask = denom > 1e-12
    c[mask] = sigma[mask] * (b[mask] - 1.0) / (denom[mask] * kappa[mask])
    return kappa, b, c


def build_cpml_profiles(prepared, dt, config=None):
    """Per-axis E (node) and H (half-node) profiles (cpml.py:380-426).

    The reference grades the high-side H samples at their cell's outer node
    (``nodes[1:]``), half a cell deeper than where they sit; at normal incidence
    that high face reflects ~6% at 12 layers / 30 cells per wavelength while the
    low face reflects ~1e-5.  It is reproduced by default for parity;
    ``config={"h_grading": "staggered"}`` grades both faces at the true H
    positions instead."""
    cfg = dict(DEFAULT_CPML_CONFIG)
    staggered = bool(config and config.get("h_grading") == "staggered")
    if config:
        cfg.update({k: v for k, v in config.items() if k in DEFAULT_CPML_CONFIG})
    scale = legacy_strength_scale(prepared.pml_strength)
    eta0 = math.sqrt(MU_0 / EPSILON_0)
    thickness = prepared.pml_thickness
    out = {}
    for axis in "xyz":
        nodes = prepared.nodes(axis)
        half = prepared.half(axis)
        lo = prepared.boundary.face_kind(axis, "low") == "pml"
        hi = prepared.boundary.face_kind(axis, "high") == "pml"
        ke, be, ce = cpml_profile_1d(nodes, nodes, thickness, dt, EPSILON_0, scale, eta0, cfg,
                                     apply_low=lo, apply_high=hi)
        kh, bh, ch = cpml_profile_1d(half, nodes, thickness, dt, MU_0, scale * (MU_0 / EPSILON_0), eta0, cfg,
                                     high_positions=None if staggered else nodes[1:], apply_low=lo, apply_high=hi)
        out[axis] = {"kappa_e": ke, "b_e": be, "c_e": ce, "kappa_h": kh, "b_h": bh, "c_h": ch}
    return out


def psi_layout(prepared, name, field_shape):
    """(axis, low_length, high_start, high_length) of one slab psi array
    (cpml.py:218-257)."""
    _, axis = PSI_SPECS[name]
    axis_name = "xyz"[axis]
    size = int(field_shape[axis])
    thickness = min(int(prepared.pml_thickness), size)
    low = thickness if prepared.boundary.face_kind(axis_name, "low") == "pml" else 0
    cap = max(size - low, 0)
    high = min(thickness if prepared.boundary.face_kind(axis_name, "high") == "pml" else 0, cap)
    return axis, low, size - high, high


def compact_shape(field_shape, layout):
    axis, low, _, high = layout
    s = list(field_shape)
    s[axis] = low + high
    return tuple(s)
